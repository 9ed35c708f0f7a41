/* A stand-in for MATLAB's mex.h with the few declarations the MEX binding in
 * INTEGRATION.md uses, so that the binding compiles and runs without MATLAB.
 * Test code only. */
#ifndef QD_TEST_MEX_H
#define QD_TEST_MEX_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct mxArray_tag mxArray;
mxArray *mxGetField(const mxArray *pm, size_t index, const char *fieldname);
double *mxGetPr(const mxArray *pa);
void *mxGetData(const mxArray *pa);
void mexErrMsgTxt(const char *msg);
mxArray *mxCreateDoubleScalar(double value);
int mexCallMATLAB(int nlhs, mxArray *plhs[], int nrhs, mxArray *prhs[], const char *fcn_name);
#ifdef __cplusplus
}
#endif
#endif
