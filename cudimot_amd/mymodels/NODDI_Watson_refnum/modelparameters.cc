#include "../NODDI_Watson/modelparameters.cc"
