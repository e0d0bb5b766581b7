// NODDI-Watson with the reference's own numerical forms (see
// modelfunctions.h): the parameters of NODDI_Watson.
#include "../NODDI_Watson/modelparameters.h"
