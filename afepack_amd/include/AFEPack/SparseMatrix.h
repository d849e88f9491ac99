// SparseMatrix.h -- drop-in CSR containers of the assemble-and-solve path:
// SparsityPattern, SparseMatrix<T>, Vector<T>, DenseMatrix<T> in the global
// namespace with the member functions the reference path uses
// (include/AFEPack/{SparsityPattern,SparseMatrix,Vector,DenseMatrix}.h).
// Host containers; a SparseMatrix<double> may be bound to a device matrix of
// the C ABI (include/afepack_amd.h), which then holds the authoritative values.
#pragma once
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <limits>
#include <stdexcept>
#include <vector>

#include "../../../include/afepack_amd.h"

const size_t invalid_size_t = std::numeric_limits<size_t>::max();

// deal.II-style pattern: square patterns keep the diagonal first in each row,
// the other columns ascending after compress() (SparsityPattern.h:16-262)
class SparsityPattern {
public:
  SparsityPattern() = default;
  SparsityPattern(size_t m, size_t n, size_t max_per_row) { reinit(m, n, max_per_row); }
  SparsityPattern(size_t m, size_t max_per_row) { reinit(m, m, max_per_row); }
  SparsityPattern(size_t m, size_t n, const std::vector<unsigned int> &len) { reinit(m, n, len); }
  SparsityPattern(size_t m, const std::vector<unsigned int> &len) { reinit(m, m, len); }

  void reinit(size_t m, size_t n, size_t max_per_row) {
    std::vector<unsigned int> len(m, (unsigned int)max_per_row);
    reinit(m, n, len);
  }
  void reinit(size_t m, size_t n, const std::vector<unsigned int> &len) {
    n_rows_ = m;
    n_cols_ = n;
    diag_first_ = (m == n);
    row_ptr_.assign(m + 1, 0);
    for (size_t i = 0; i < m; i++) row_ptr_[i + 1] = row_ptr_[i] + len[i];
    col_.assign(row_ptr_[m], invalid_size_t);
    if (diag_first_)
      for (size_t i = 0; i < m; i++)
        if (row_ptr_[i + 1] > row_ptr_[i]) col_[row_ptr_[i]] = i;
    compressed_ = false;
  }
  // adopt an already compressed pattern (row_ptr, col)
  void assign_compressed(size_t m, size_t n, std::vector<size_t> rp, std::vector<size_t> col) {
    n_rows_ = m;
    n_cols_ = n;
    diag_first_ = (m == n);
    row_ptr_ = std::move(rp);
    col_ = std::move(col);
    compressed_ = true;
  }
  void add(size_t i, size_t j) {
    for (size_t k = row_ptr_[i]; k < row_ptr_[i + 1]; k++) {
      if (col_[k] == j) return;
      if (col_[k] == invalid_size_t) {
        col_[k] = j;
        return;
      }
    }
  }
  void compress() {
    if (compressed_) return;
    size_t w = 0, start = 0;
    for (size_t i = 0; i < n_rows_; i++) {
      const size_t end = row_ptr_[i + 1];
      size_t cnt = 0;
      for (size_t k = start; k < end && col_[k] != invalid_size_t; k++) col_[w + cnt++] = col_[k];
      std::sort(col_.begin() + w + (diag_first_ && cnt ? 1 : 0), col_.begin() + w + cnt);
      row_ptr_[i] = w;
      w += cnt;
      start = end;
    }
    row_ptr_[n_rows_] = w;
    col_.resize(w);
    compressed_ = true;
  }
  size_t max_entries_per_row() const {
    size_t m = 0;
    for (size_t i = 1; i < n_rows_; i++) m = std::max(m, row_ptr_[i] - row_ptr_[i - 1]);
    return m;
  }
  size_t n_rows() const { return n_rows_; }
  size_t n_cols() const { return n_cols_; }
  size_t n_nnz() const { return row_ptr_.empty() ? 0 : row_ptr_[n_rows_]; }
  size_t n_nonzero_elements() const { return n_nnz(); }
  size_t row_length(size_t r) const { return row_ptr_[r + 1] - row_ptr_[r]; }
  bool is_compressed() const { return compressed_; }
  const std::vector<size_t> &row_ptr() const { return row_ptr_; }
  const std::vector<size_t> &col_indices() const { return col_; }
  const size_t *get_rowstart_indices() const { return row_ptr_.data(); }
  const size_t *get_column_numbers() const { return col_.data(); }

private:
  size_t n_rows_ = 0, n_cols_ = 0;
  bool diag_first_ = false, compressed_ = false;
  std::vector<size_t> row_ptr_, col_;
};

template <class T> class Vector {
public:
  Vector() = default;
  explicit Vector(size_t n) : d_(n, T(0)) {}
  void reinit(size_t n) { d_.assign(n, T(0)); }
  size_t size() const { return d_.size(); }
  T &operator()(size_t i) { return d_[i]; }
  T operator()(size_t i) const { return d_[i]; }
  T &operator[](size_t i) { return d_[i]; }
  T operator[](size_t i) const { return d_[i]; }
  Vector &operator=(T v) {
    std::fill(d_.begin(), d_.end(), v);
    return *this;
  }
  void add(T a, const Vector &v) {
    if (v.size() != size()) throw std::runtime_error("vector size mismatch");
    for (size_t i = 0; i < size(); i++) d_[i] += a * v.d_[i];
  }
  Vector &operator+=(const Vector &v) {
    add(T(1), v);
    return *this;
  }
  Vector &operator-=(const Vector &v) {
    add(T(-1), v);
    return *this;
  }
  T l1_norm() const {
    T s = 0;
    for (T v : d_) s += std::fabs(v);
    return s;
  }
  T l2_norm() const {
    T s = 0;
    for (T v : d_) s += v * v;
    return std::sqrt(s);
  }
  T *begin() { return d_.data(); }
  const T *begin() const { return d_.data(); }
  T *data() { return d_.data(); }
  const T *data() const { return d_.data(); }

private:
  std::vector<T> d_;
};

template <class T> class DenseMatrix {
public:
  DenseMatrix() = default;
  DenseMatrix(size_t m, size_t n) { reinit(m, n); }
  void reinit(size_t m, size_t n) {
    m_ = m;
    n_ = n;
    d_.assign(m * n, T(0));
  }
  T &operator()(size_t i, size_t j) { return d_[i * n_ + j]; }
  T operator()(size_t i, size_t j) const { return d_[i * n_ + j]; }
  size_t n_rows() const { return m_; }
  size_t n_cols() const { return n_; }
  size_t m() const { return m_; }
  size_t n() const { return n_; }
  const T *data() const { return d_.data(); }

private:
  size_t m_ = 0, n_ = 0;
  std::vector<T> d_;
};

// CSR matrix over a SparsityPattern (SparseMatrix.h:14-144).  When bound to a
// device matrix (bind_device) the two copies are kept coherent lazily: device
// operations mark the host copy stale and the next host accessor copies the
// values back; host-side writes mark the device copy stale and the next device
// use pushes them (push_to_device).
template <class T> class SparseMatrix {
public:
  SparseMatrix() = default;
  explicit SparseMatrix(SparsityPattern &sp) { reinit(sp); }
  virtual ~SparseMatrix() { release_device(); }
  SparseMatrix(const SparseMatrix &) = delete;
  SparseMatrix &operator=(const SparseMatrix &) = delete;

  void reinit(SparsityPattern &sp) {
    release_device();
    pattern_ = &sp;
    val_.assign(sp.n_nnz(), T(0));
  }
  void add(size_t i, size_t j, T v) {
    ensure_host();
    const size_t *rp = pattern_->get_rowstart_indices(), *col = pattern_->get_column_numbers();
    for (size_t k = rp[i]; k < rp[i + 1]; k++)
      if (col[k] == j) {
        val_[k] += v;
        host_dirty_ = true;
        return;
      }
  }
  T el(size_t i, size_t j) const {
    ensure_host();
    const size_t *rp = pattern_->get_rowstart_indices(), *col = pattern_->get_column_numbers();
    for (size_t k = rp[i]; k < rp[i + 1]; k++)
      if (col[k] == j) return val_[k];
    return 0;
  }
  T diag_element(size_t i) const {
    ensure_host();
    return val_[pattern_->get_rowstart_indices()[i]];
  }
  T global_entry(size_t k) const {
    ensure_host();
    return val_[k];
  }
  T &global_entry(size_t k) {
    ensure_host();
    host_dirty_ = true;
    return val_[k];
  }
  size_t m() const { return pattern_ ? pattern_->n_rows() : 0; }
  size_t n() const { return pattern_ ? pattern_->n_cols() : 0; }
  size_t n_nonzero_elements() const { return pattern_->n_nnz(); }
  const SparsityPattern &get_sparsity_pattern() const { return *pattern_; }
  void vmult(Vector<T> &dst, const Vector<T> &src) const {
    ensure_host();
    const size_t *rp = pattern_->get_rowstart_indices(), *col = pattern_->get_column_numbers();
    dst = T(0);
    for (size_t i = 0; i < m(); i++)
      for (size_t k = rp[i]; k < rp[i + 1]; k++) dst(i) += val_[k] * src(col[k]);
  }
  void Tvmult(Vector<T> &dst, const Vector<T> &src) const {
    ensure_host();
    const size_t *rp = pattern_->get_rowstart_indices(), *col = pattern_->get_column_numbers();
    dst = T(0);
    for (size_t i = 0; i < m(); i++)
      for (size_t k = rp[i]; k < rp[i + 1]; k++) dst(col[k]) += val_[k] * src(i);
  }
  const std::vector<T> &values() const {
    ensure_host();
    return val_;
  }

  // ---- device binding ---------------------------------------------------
  afe_matrix *device_matrix() const { return dev_; }
  void bind_device(afe_matrix *a, bool owned) {
    release_device();
    dev_ = a;
    dev_owned_ = owned;
  }
  // the device copy was changed by a device operation
  void mark_device_newer() {
    host_stale_ = true;
    host_dirty_ = false;
  }
  void ensure_host() const {
    if (dev_ && host_stale_) {
      if (!val_.empty()) check(afe_matrix_values(dev_, val_.data()));
      host_stale_ = false;
    }
  }
  // host -> device when host-side writes happened since the last sync
  void push_to_device() const {
    if (dev_ && host_dirty_) {
      check(afe_matrix_set_values(dev_, val_.data()));
      host_dirty_ = false;
    }
  }
  static void check(int rc) {
    if (rc != AFE_OK) throw std::runtime_error(afe_last_error());
  }

protected:
  void release_device() {
    if (dev_ && dev_owned_) afe_matrix_destroy(dev_);
    dev_ = nullptr;
    host_stale_ = host_dirty_ = false;
  }
  SparsityPattern *pattern_ = nullptr;
  mutable std::vector<T> val_;
  afe_matrix *dev_ = nullptr;
  bool dev_owned_ = false;
  mutable bool host_stale_ = false, host_dirty_ = false;
};
