"""relxill_amd -- MI355X (gfx950) implementation of relxill's relativistic
smearing path behind the reference's XSPEC local-model C ABI.

The product is the C-ABI library ``librelxill_gpu.so`` (C++ host + HIP
kernels).  This package is a thin Python mirror of that ABI (ctypes) used by
tests, benchmarks and interactive work; it has no compute path of its own and
raises ImportError when the HIP library is missing.  Names are resolved lazily
so that ``python -m relxill_amd.build`` works before the library exists.
"""

_LIB_NAMES = {
    "FP32", "FP64", "MODEL_RELCONV", "MODEL_RELCONVLP", "MODEL_RELLINE", "MODEL_RELLINELP",
    "MODEL_RELXILL_CONV", "MODEL_RELXILL", "MODEL_RELXILLLP", "DEFAULTS", "N_PARAMS", "Batch", "RxError", "eval_model", "lib",
    "lmodrelconv", "lmodrelconvlp", "lmodrelline", "lmodrellinelp", "lmodrelxill", "lmodrelxilllp", "model_params",
    "nthcomp", "MODEL_RELXILLCP", "MODEL_RELXILLLPCP", "lmodrelxilldensnthcomp", "lmodrelxilllpdensnthcomp",
    "MODEL_RELXILLNS", "lmodrelxillns", "xillnstable_synthetic",
    "MODEL_XILLVER", "MODEL_XILLVERCP", "MODEL_XILLVERNS", "lmodxillver", "lmodxillverdensnthcomp", "lmodxillverns",
    "xilltable_synthetic", "xillcptable_synthetic", "xillcptable_synthetic_dens", "xilltable_view",
    "rradtable_synthetic", "rradtable_view",
    "reflset_create", "set_conv_grid", "tables_synthetic", "lptable_synthetic", "tables_view",
    "Server", "RX_OK", "RX_ERR_PARAM", "RX_ERR_UNSUPPORTED", "RX_ERR_TABLE", "RX_ERR_INPUT",
    "RX_ERR_DEVICE", "RX_ERR_ARG",
}

__all__ = sorted(_LIB_NAMES)


def __getattr__(name):
    if name in _LIB_NAMES:
        import importlib
        return getattr(importlib.import_module(".capi", __name__), name)
    raise AttributeError(name)
