"""MI355X-native closed-shell coupled-cluster hot path (CCSD residual and
iteration, (T)), the host-side mirror of Dalton's CC driver interface over
the C ABI of libccgpu.so (include/ccgpu.h)."""
from .cc import (CCBatch, CCContext, df_factor_ao, df_factor_ao_cols, dgemm, FP32, mfma_ceiling, FP64, MIXED, npack, pack_t2, rsp_filename,  # noqa: F401
                 rsp_read, rsp_write, unpack_t2)
