// Event-queue (batched) transport form.
#include "omc_capi_internal.hpp"

template <class R>
int omc_run_dosxyz_event(omc_ctx *c, PhantomDev *pd, const omc_dosxyz_source_desc *src, int64_t nper, int nbatch,
                         const omc_run_opts *o, double *sum, double *sum2, omc_run_stats *st) {
    (void)c; (void)pd; (void)src; (void)nper; (void)nbatch; (void)o; (void)sum; (void)sum2; (void)st;
    return omc_seterr(OMC_ERR_ARG, "event form not available in this build");
}

template int omc_run_dosxyz_event<float>(omc_ctx *, PhantomDev *, const omc_dosxyz_source_desc *, int64_t, int,
                                         const omc_run_opts *, double *, double *, omc_run_stats *);
template int omc_run_dosxyz_event<double>(omc_ctx *, PhantomDev *, const omc_dosxyz_source_desc *, int64_t, int,
                                          const omc_run_opts *, double *, double *, omc_run_stats *);

extern "C" int omc_run_dij(omc_ctx *c, int phantom_id, const omc_spectrum_desc *spec, const omc_beamlets_desc *bix,
                           int nbatch, double rel_threshold, const omc_run_opts *opts, omc_csc *dij, omc_csc *var,
                           omc_run_stats *stats) {
    (void)c; (void)phantom_id; (void)spec; (void)bix; (void)nbatch; (void)rel_threshold; (void)opts; (void)dij; (void)var; (void)stats;
    return omc_seterr(OMC_ERR_ARG, "Dij not available in this build");
}
extern "C" void omc_csc_free(omc_csc *m) {
    if (!m) return;
    free(m->jc); free(m->ir); free(m->val);
    m->jc = m->ir = nullptr; m->val = nullptr; m->nnz = 0;
}
extern "C" int omc_server_create(omc_ctx *, int64_t, int, omc_server **) { return omc_seterr(OMC_ERR_ARG, "serving not available in this build"); }
extern "C" int omc_server_submit(omc_server *, const omc_request *, int64_t *) { return OMC_ERR_ARG; }
extern "C" int omc_server_poll(omc_server *, int64_t, int *) { return OMC_ERR_ARG; }
extern "C" int omc_server_wait(omc_server *, int64_t, omc_result *) { return OMC_ERR_ARG; }
extern "C" void omc_result_free(omc_result *) {}
extern "C" void omc_server_destroy(omc_server *) {}
