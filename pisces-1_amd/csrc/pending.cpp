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
int pisces_server_create(pisces_server **, int, size_t, int, int, int, int) { return PISCES_ERR_UNSUPPORTED; }
void pisces_server_destroy(pisces_server *) {}
long long pisces_server_submit(pisces_server *, const pisces_request *) { return PISCES_ERR_UNSUPPORTED; }
int pisces_server_step(pisces_server *) { return PISCES_ERR_UNSUPPORTED; }
int pisces_server_poll(pisces_server *, pisces_result *, int) { return PISCES_ERR_UNSUPPORTED; }
size_t pisces_server_footprint(const pisces_server *, const pisces_request *) { return 0; }
int pisces_server_vectors(pisces_server *, long long, double *) { return PISCES_ERR_UNSUPPORTED; }
}
