// placeholder, replaced below
#include "ngp_common.h"
