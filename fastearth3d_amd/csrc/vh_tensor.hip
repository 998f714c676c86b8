// vh_tensor.hip -- placeholder until the lateral-viscosity path lands.
#include "vh_internal.h"
