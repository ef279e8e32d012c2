// vmx_f64.hip — FP64 instantiation of the kernels.  Built with -ffp-contract=off: every
// product and sum rounds separately, in the order written, as in the reference CPU code.
#define VMX_REAL double
#include "vmx_ops.inc"
