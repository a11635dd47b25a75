// Work-queue kernels of mixed-length plans (k_queue, k_queue_setup) and
// their launchers: kernels.hip compiled with GOFFT_QUEUE_TU defined, in a
// translation unit of their own so that the two compile in parallel.
#define GOFFT_QUEUE_TU 1
#include "kernels.hip"
