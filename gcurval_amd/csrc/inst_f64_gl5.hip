// Kernel instantiations for double arithmetic with 5-point Gauss-Legendre.
#include "launch_impl.h"

template struct gcurval::rt::Launch<double, 5>;
