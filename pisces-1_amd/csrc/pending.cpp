// pending.cpp -- entry points whose implementation is added later in the build
// order; they fail loudly instead of silently doing nothing.
#include "context.hpp"

extern "C" {
int pisces_davidson_work_size(int, int, int, int) { return PISCES_ERR_UNSUPPORTED; }
int pisces_davidson(pisces_ctx *, int, int, int, int, int, int, int, double *, int *, double *,
                    double *, const double *, double *, int *)
{
    return PISCES_ERR_UNSUPPORTED;
}
}
