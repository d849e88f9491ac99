// C entry points over the reference's own SparsityPattern class
// (include/AFEPack/SparsityPattern.h, header-only and self-contained apart
// from <vector>).  Built by oracle/Makefile into oracle/_ref/ only when the
// reference tree is present; used by the oracle-pinning tests to check the
// oracle's pattern mechanics (reinit/add/compress) against the real class.
#include <cstddef>
#include <cstdint>
#include <vector>
#include "SparsityPattern.h"

extern "C" {
void *refsp_new_max(int64_t m, int64_t n, int64_t maxlen) {
  return new SparsityPattern((size_t)m, (size_t)n, (size_t)maxlen);
}
void *refsp_new_lengths(int64_t m, int64_t n, const int64_t *len) {
  std::vector<unsigned int> l(len, len + m);
  return new SparsityPattern((size_t)m, (size_t)n, l);
}
void refsp_add(void *p, int64_t i, int64_t j) {
  static_cast<SparsityPattern *>(p)->add((size_t)i, (size_t)j);
}
void refsp_compress(void *p) { static_cast<SparsityPattern *>(p)->compress(); }
int64_t refsp_nnz(void *p) {
  auto *sp = static_cast<SparsityPattern *>(p);
  return (int64_t)sp->get_rowstart_indices()[sp->n_rows()];
}
int64_t refsp_max_entries_per_row(void *p) {
  return (int64_t) static_cast<SparsityPattern *>(p)->max_entries_per_row();
}
void refsp_get(void *p, int64_t *rp, int64_t *col) {
  auto *sp = static_cast<SparsityPattern *>(p);
  size_t m = sp->n_rows();
  const size_t *r = sp->get_rowstart_indices();
  const size_t *c = sp->get_column_numbers();
  for (size_t i = 0; i <= m; i++) rp[i] = (int64_t)r[i];
  for (size_t k = 0; k < r[m]; k++) col[k] = (int64_t)c[k];
}
void refsp_free(void *p) { delete static_cast<SparsityPattern *>(p); }
}
