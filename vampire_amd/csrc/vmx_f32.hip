// vmx_f32.hip — FP32 instantiation of the kernels (FMA contraction allowed).
#define VMX_REAL float
#include "vmx_ops.inc"
