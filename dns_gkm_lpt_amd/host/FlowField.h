// FlowField.h -- drop-in for the reference's src/FlowField.h (FlowField.cpp:9-337).
#ifndef FLOWFIELDHEADERDEF
#define FLOWFIELDHEADERDEF

#include "param.h"

void GetParams();
void Initialize_DHIT(ptype *W[5]);
void PeriodicBC(ptype *W[5]);
void segment();
void segmentFill(ptype *W[5], ptype *W0[5]);
void PrintParams();

#endif
