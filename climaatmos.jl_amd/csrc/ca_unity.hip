// ca_unity.hip -- single translation unit of libcadycore (the GLL constant
// tables live in __constant__ memory shared by all kernels).
#include "ca_column.hip"
#include "ca_newton_pcr.hip"
#include "ca_spectral.hip"
#include "ca_dss.hip"
#include "ca_ensemble.hip"
#include "ca_adjoint.hip"
#include "ca_step_adjoint.hip"
#include "ca_dss_table.cpp"
#include "ca_api.hip"
#include "ca_ensemble_api.hip"
#include "ca_step_adjoint_api.hip"
#include "ca_hooks_fv.hip"
