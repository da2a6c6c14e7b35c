// Continuous-batching request server: placeholder translation unit.
#include "phx_serve.h"
