// ps_internal.h -- declarations shared between the library's source files;
// not part of the C ABI.
#pragma once

#include <string>

#include "periodsearch.h"

// Integer triangulation of the Gaussian sphere (period_search/trifac.c:16-79):
// ifp[F][3] with 1-based vertex ids, F = 8 nrows^2.  Defined in ps_host.cpp.
void psamd_triangulate(int nrows, int *ifp);

// Engine internals used by the shape-setup kernels (ps_shape.hip); defined
// in ps_engine.hip.  The stream is the engine's hipStream_t.
void *psamd_engine_stream(ps_engine *e);
int psamd_engine_device(const ps_engine *e);
void psamd_engine_error(ps_engine *e, const std::string &msg);
