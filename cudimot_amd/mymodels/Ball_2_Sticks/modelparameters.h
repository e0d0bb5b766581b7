#ifndef MODEL_H_INCLUDED
#define MODEL_H_INCLUDED
// Parameters: S0, d, f1, th1, ph1, f2, th2, ph2.  CFP: bvecs (3), bvals (1).
// MyType selects the precision the engine evaluates this model's forward
// signal and Jacobian in (float: FP32 on the GPU).
typedef float MyType;
#define NPARAMS 8
#define NCFP 2
// Levenberg-Marquardt Jacobian in double (engine option): near-degenerate
// directions make single-precision central differences noise-dominated.
#define CDM_LM_JACOBIAN_DOUBLE
#define NFIXP 0

struct MODEL {
  static int CFP_size[NCFP];
  static int FixP_size[NFIXP];
};
#endif
