// Kernel instantiations for float arithmetic with 3-point Gauss-Legendre.
#include "launch_impl.h"

template struct gcurval::rt::Launch<float, 3>;
