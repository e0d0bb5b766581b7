#ifndef MODEL_H_INCLUDED
#define MODEL_H_INCLUDED
// Parameters: fiso, fintra, kappa, beta, th, ph, psi.  CFP: bvecs (3), bvals (1).  FixP: S0.
// MyType selects the precision the engine evaluates this model's forward
// signal and Jacobian in (float: FP32 on the GPU).
typedef float MyType;
#define NPARAMS 7
#define NCFP 2
// Levenberg-Marquardt Jacobian in double (engine option): near-degenerate
// directions make single-precision central differences noise-dominated.
#define CDM_LM_JACOBIAN_DOUBLE
#define NFIXP 1

struct MODEL {
  static int CFP_size[NCFP];
  static int FixP_size[NFIXP];
};
#endif
