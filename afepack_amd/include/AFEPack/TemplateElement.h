// TemplateElement.h -- part of the drop-in FE space interface (FEMSpace.h).
#pragma once
#include "FEMSpace.h"
