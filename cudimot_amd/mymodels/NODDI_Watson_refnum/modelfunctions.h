// NODDI-Watson, reference-parity numerics: the NODDI_Watson plug-in with
// the two numerical forms it changes restored to the reference's
// (NODDI_Watson/WatsonFunctions.h:91-194, LegendreFunctions.h:6-57):
//  * the closed-form spherical-harmonic coefficients of the Watson density
//    for 0.4 < kappa <= 3 (NODDI_Watson takes them by quadrature there, since
//    the closed forms cancel against the ~2e-7 Dawson approximation);
//  * erf in single precision (of the float-rounded argument, correctly
//    rounded) in the Legendre-Gaussian integrals (NODDI_Watson uses double
//    erf).
// Everything else, caches and precision split included, is NODDI_Watson's.
// Use it where GridSearch argmins or fits must follow the reference's
// arithmetic; NODDI_Watson is the more accurate model.
#define CDM_WATSON_REFERENCE_FORMS 1
#include "../NODDI_Watson/modelfunctions.h"
