// SzSymmetry.h: the header name under which the reference declares
// Symmetry::SzSymmetry (SzSymmetry.h:19-234); this build keeps that class in
// Symmetry.h.
#ifndef EDLIB_AMD_SZSYMMETRY_H
#define EDLIB_AMD_SZSYMMETRY_H

#include "Symmetry.h"

#endif
