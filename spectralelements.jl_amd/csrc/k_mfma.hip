// MFMA contraction kernels (FP16/BF16 in, FP32 accumulate) for nr == ns == 16.
#include "sem_device.h"
