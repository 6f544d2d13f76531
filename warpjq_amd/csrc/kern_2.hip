// k_main / k_emit instantiated for a 2-entry slot table.
#include "kernels.h"

WJQ_DEFINE_LAUNCHERS(2)
