// dnsgkm_host.h -- the solver context shared by the drop-in entry points.
#ifndef DNSGKM_HOST_H
#define DNSGKM_HOST_H

#include "../../include/dnsgkm.h"

// The context of the run described by the param.h globals (created on first
// use after GetParams); every entry point of the drop-in layer uses it.
gkm_ctx *dnsgkm_context();
// Exit with the context's error message when rc is not GKM_OK.
void dnsgkm_check(int rc, const char *what);

#endif
