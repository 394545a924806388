// sfh_batched.hip -- batched fg! over many coefficient vectors (placeholder).
#include "sfh_internal.h"

using namespace sfh;

extern "C" {

int sfh_fg_batched_device(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                          const double* d_X, int64_t B, int precision, double eps,
                          double* d_negLL, double* d_G)
{
    (void)tpl; (void)data; (void)d_X; (void)B; (void)precision; (void)eps; (void)d_negLL; (void)d_G;
    if (!ctx) return SFH_ERR_ARGUMENT;
    return set_error(ctx, SFH_ERR_ARGUMENT, "batched path not built yet");
}

int sfh_fg_batched(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data, const double* X,
                   int64_t B, int precision, double eps, double* negLL, double* G)
{
    return sfh_fg_batched_device(ctx, tpl, data, X, B, precision, eps, negLL, G);
}

}
