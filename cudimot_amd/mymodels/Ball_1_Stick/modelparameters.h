#ifndef MODEL_H_INCLUDED
#define MODEL_H_INCLUDED
// Parameters: S0, d, f, th, ph.  CFP: bvecs (3), bvals (1).
// MyType selects the precision the engine evaluates this model's forward
// signal and Jacobian in (float: FP32 on the GPU).
typedef float MyType;
#define NPARAMS 5
#define NCFP 2
#define NFIXP 0

struct MODEL {
  static int CFP_size[NCFP];
  static int FixP_size[NFIXP];
};
#endif
