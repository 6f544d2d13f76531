// k_main / k_emit instantiated for a 1-entry slot table.
#include "kernels.h"

WJQ_DEFINE_LAUNCHERS(1)
