"""MI355X-native voxelwise diffusion-MRI model fitting (cudimot's fitting path).

The engine is a HIP library (``lib/libcudimot.so``, built by
``python -m cudimot_amd.build``) with a C-ABI (``include/cudimot_amd.h``);
``cudimot_amd.capi`` binds it for Python callers.
"""
from .capi import Model, CdmError, list_models, sync  # noqa: F401

GRIDSEARCH, LEVMAR, MCMC, PREDICTED, BIC_AIC = range(5)
