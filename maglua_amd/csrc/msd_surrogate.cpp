// Learned dipole-field surrogate of libmsd.so: weights, training, inference and
// its checks against the exact field.
#include "msd_host.h"

static int surrogate_dims_ok(int nx, int ny, int nz)
{
    if (nx < 32 || nx % 32) return fail(MSD_EUNSUPPORTED, "Surrogate.new: nx=%d must be a multiple of 32", nx);
    if (ny < 1 || nz < 1) return fail(MSD_EINVAL, "Surrogate.new: bad size");
    if ((double)nx * ny * nz > (double)(1u << 27)) return fail(MSD_EUNSUPPORTED, "Surrogate.new: more than 2^27 sites");
    return MSD_OK;
}

extern "C" int msd_surrogate_set_open(msd_surrogate* s, int open)
{
    NONNULL(s);
    s->cd.open = open ? 1 : 0;
    s->plan_walks();
    return MSD_OK;
}

extern "C" int msd_surrogate_create(int nx, int ny, int nz, msd_surrogate** out)
{
    NONNULL(out);
    *out = nullptr;
    if (int rc = surrogate_dims_ok(nx, ny, nz)) return rc;
    if (int rc = ensure_init()) return rc;
    std::unique_ptr<msd_surrogate> s(new msd_surrogate);
    s->d = Dims{nx, ny, nz, nx * ny * nz};
    CnnDims& c = s->cd;
    c.nx = nx; c.ny = ny; c.nz = nz; c.n = s->d.n;
    c.tiles_x = nx / 32; c.tiles_y = (ny + 7) / 8; c.ntiles = c.tiles_x * c.tiles_y * nz;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, G.device));
    s->cus = prop.multiProcessorCount;
    s->grid = std::min(c.ntiles, 2 * s->cus);
    s->plan_walks();
    const size_t n = s->d.n, W3 = (size_t)3 * msd_surrogate::WSZ;
    for (float** p : {&s->W, &s->Mo, &s->Vo, &s->gW})
        if (int rc = dalloc(p, W3)) return rc;
    if (int rc = dalloc(&s->Af, (size_t)3 * msd_surrogate::PACK)) return rc;
    if (int rc = dalloc(&s->At, (size_t)3 * msd_surrogate::PACK)) return rc;
    for (uint16_t** p : {&s->X1, &s->X2, &s->G1, &s->G2})
        if (int rc = dalloc(p, 16 * n)) return rc;
    for (uint16_t** p : {&s->X0, &s->G3})
        if (int rc = dalloc(p, 4 * n)) return rc;
    if (int rc = dalloc(&s->zeros, 8)) return rc;
    for (float** p : {&s->Y3, &s->T, &s->D})
        if (int rc = dalloc(p, 3 * n)) return rc;
    if (int rc = dalloc(&s->partial, (size_t)s->grid * msd_surrogate::WSZ)) return rc;
    s->nlpart = (int)((n + 255) / 256);
    if (int rc = dalloc(&s->lpart, (size_t)s->nlpart)) return rc;
    hipError_t e = hipSuccess;
    for (float* p : {s->W, s->Mo, s->Vo, s->gW})
        if (e == hipSuccess) e = launch_fill(p, 0.0f, W3, G.stream);
    if (e == hipSuccess) e = launch_fill(reinterpret_cast<float*>(s->zeros), 0.0f, 4, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.new: %s", hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(G.stream));
    *out = s.release();
    return MSD_OK;
}

extern "C" int msd_surrogate_destroy(msd_surrogate* s) { delete s; return MSD_OK; }

extern "C" int msd_surrogate_set_weights(msd_surrogate* s, int layer, const double* w)
{
    NONNULL(s); NONNULL(w);
    if (layer < 0 || layer > 2) return fail(MSD_EINVAL, "Surrogate.setWeights: layer %d", layer);
    const int co = msd_surrogate::cout(layer), ci = msd_surrogate::cin(layer);
    std::vector<float> h(msd_surrogate::WSZ, 0.0f);
    for (int t = 0; t < msd_surrogate::NT; ++t)
        for (int o = 0; o < co; ++o)
            for (int i = 0; i < ci; ++i) h[(t * 16 + o) * 16 + i] = (float)w[(t * co + o) * ci + i];
    HIPCHK(hipMemcpyAsync(s->W + (size_t)layer * msd_surrogate::WSZ, h.data(), h.size() * sizeof(float),
                          hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    s->packed = false;
    return MSD_OK;
}

static int surrogate_get(const float* dsrc, int layer, double* w)
{
    const int co = msd_surrogate::cout(layer), ci = msd_surrogate::cin(layer);
    std::vector<float> h(msd_surrogate::WSZ);
    HIPCHK(hipMemcpyAsync(h.data(), dsrc + (size_t)layer * msd_surrogate::WSZ, h.size() * sizeof(float),
                          hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    for (int t = 0; t < msd_surrogate::NT; ++t)
        for (int o = 0; o < co; ++o)
            for (int i = 0; i < ci; ++i) w[(t * co + o) * ci + i] = h[(t * 16 + o) * 16 + i];
    return MSD_OK;
}

extern "C" int msd_surrogate_get_weights(const msd_surrogate* s, int layer, double* w)
{
    NONNULL(s); NONNULL(w);
    if (layer < 0 || layer > 2) return fail(MSD_EINVAL, "Surrogate.weights: layer %d", layer);
    return surrogate_get(s->W, layer, w);
}

extern "C" int msd_surrogate_get_grads(const msd_surrogate* s, int layer, double* g)
{
    NONNULL(s); NONNULL(g);
    if (layer < 0 || layer > 2) return fail(MSD_EINVAL, "Surrogate.grads: layer %d", layer);
    return surrogate_get(s->gW, layer, g);
}

extern "C" int msd_surrogate_set_scales(msd_surrogate* s, double s_in, double s_out, double sigma2)
{
    NONNULL(s);
    if (!(s_in > 0) || !(s_out > 0) || !(sigma2 > 0)) return fail(MSD_EINVAL, "Surrogate.setScales: scales must be > 0");
    s->s_in = s_in; s->s_out = s_out; s->sigma2 = sigma2;
    return MSD_OK;
}

extern "C" int msd_surrogate_get_scales(const msd_surrogate* s, double* s_in, double* s_out, double* sigma2)
{
    NONNULL(s); NONNULL(s_in); NONNULL(s_out); NONNULL(sigma2);
    *s_in = s->s_in; *s_out = s->s_out; *sigma2 = s->sigma2;
    return MSD_OK;
}

extern "C" int msd_surrogate_set_input(msd_surrogate* s, const double* dm)
{
    NONNULL(s); NONNULL(dm);
    if (int rc = upload_f32(s->D, dm, (size_t)3 * s->d.n)) return rc;
    hipError_t e = launch_cnn_input(s->d, s->D, nullptr, (float)s->s_in, s->X0, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.setInput: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_surrogate_set_input_systems(msd_surrogate* s, const msd_system* m, const msd_system* m_ref)
{
    NONNULL(s); NONNULL(m); NONNULL(m_ref);
    if (!same_dims(s->d, m->d.nx, m->d.ny, m->d.nz) || !same_dims(s->d, m_ref->d.nx, m_ref->d.ny, m_ref->d.nz))
        return fail(MSD_EMISMATCH, "Surrogate.setInput: SpinSystem size differs");
    hipError_t e = launch_cnn_input(s->d, m->S, m_ref->S, (float)s->s_in, s->X0, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.setInput: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_surrogate_set_target(msd_surrogate* s, const double* t)
{
    NONNULL(s); NONNULL(t);
    return upload_f32(s->T, t, (size_t)3 * s->d.n);
}


static int surrogate_set_target(msd_surrogate* s, ExactField f, const msd_system* m, const msd_system* m_ref)
{
    if (!same_dims(s->d, m->d.nx, m->d.ny, m->d.nz) || !same_dims(s->d, m_ref->d.nx, m_ref->d.ny, m_ref->d.nz) ||
        !f.fits(s->d))
        return fail(MSD_EMISMATCH, "Surrogate.setTarget: SpinSystem or field operator size differs");
    hipError_t e = launch_cnn_sub((size_t)3 * s->d.n, m->S, m_ref->S, s->D, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.setTarget: %s", hipGetErrorString(e));
    return f.of(s->D, s->T);   // the field is linear in the spins
}

extern "C" int msd_surrogate_set_target_dipole(msd_surrogate* s, msd_dipole* dip, const msd_system* m,
                                               const msd_system* m_ref)
{
    NONNULL(s); NONNULL(dip); NONNULL(m); NONNULL(m_ref);
    return surrogate_set_target(s, ExactField{dip, nullptr}, m, m_ref);
}

extern "C" int msd_surrogate_set_target_longrange3d(msd_surrogate* s, msd_longrange3d* lr, const msd_system* m,
                                                    const msd_system* m_ref)
{
    NONNULL(s); NONNULL(lr); NONNULL(m); NONNULL(m_ref);
    return surrogate_set_target(s, ExactField{nullptr, lr}, m, m_ref);
}

extern "C" int msd_surrogate_forward(msd_surrogate* s, double* out)
{
    NONNULL(s);
    if (int rc = s->forward(s->Y3, nullptr, G.stream)) return rc;
    if (out) return download_f32(out, s->Y3, (size_t)3 * s->d.n);
    return MSD_OK;
}

// forward, loss, back-propagation and the three weight gradients
static int surrogate_loss_grads(msd_surrogate* s, double* loss)
{
    const size_t n = s->d.n;
    if (int rc = s->forward(s->Y3, nullptr, G.stream)) return rc;
    const float gscale = (float)(2.0 / (3.0 * (double)n * s->sigma2));
    const int P = msd_surrogate::PACK, WS = msd_surrogate::WSZ;
    // the forward output already carries s_out: dL/dY3 = s_out dL/d(dH)
    hipError_t e = launch_cnn_loss(s->d, s->Y3, s->T, 1.0f, gscale * (float)s->s_out, s->G3, s->lpart, G.stream);
    if (e == hipSuccess) e = s->rows(CV_IN3, s->G3, s->At + 2 * P, s->G2, nullptr, 1.0f, G.stream);
    if (e == hipSuccess) e = s->rows(CV_F16, s->G2, s->At + P, s->G1, nullptr, 1.0f, G.stream);
    if (e == hipSuccess) e = launch_cnn_wgrad(s->cd, s->grid, s->X2, 16, s->G3, 4, s->partial, s->gW + 2 * WS, G.stream);
    if (e == hipSuccess) e = launch_cnn_wgrad(s->cd, s->grid, s->X1, 16, s->G2, 16, s->partial, s->gW + WS, G.stream);
    if (e == hipSuccess) e = launch_cnn_wgrad(s->cd, s->grid, s->X0, 4, s->G1, 16, s->partial, s->gW, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "surrogate backward: %s", hipGetErrorString(e));
    if (loss) {
        std::vector<double> h(s->nlpart);
        HIPCHK(hipMemcpyAsync(h.data(), s->lpart, sizeof(double) * h.size(), hipMemcpyDeviceToHost, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
        double acc = 0.0;
        for (double v : h) acc += v;
        *loss = acc / (3.0 * (double)n * s->sigma2);
    }
    return MSD_OK;
}

extern "C" int msd_surrogate_loss_grads(msd_surrogate* s, double* loss)
{
    NONNULL(s);
    return surrogate_loss_grads(s, loss);
}

extern "C" int msd_surrogate_adam(msd_surrogate* s, double lr)
{
    NONNULL(s);
    if (!(lr > 0)) return fail(MSD_EINVAL, "Surrogate.adam: lr must be > 0");
    const double b1 = 0.9, b2 = 0.999;
    s->adam_t++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, s->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, s->adam_t)));
    hipError_t e = launch_cnn_adam(s->W, s->gW, s->Mo, s->Vo, 3 * msd_surrogate::WSZ, (float)lr, (float)b1, (float)b2,
                                   c1, c2, 1e-8f, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.adam: %s", hipGetErrorString(e));
    s->packed = false;
    return MSD_OK;
}

extern "C" int msd_surrogate_reset_optimizer(msd_surrogate* s)
{
    NONNULL(s);
    hipError_t e = launch_fill(s->Mo, 0.0f, (size_t)3 * msd_surrogate::WSZ, G.stream);
    if (e == hipSuccess) e = launch_fill(s->Vo, 0.0f, (size_t)3 * msd_surrogate::WSZ, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.resetOptimizer: %s", hipGetErrorString(e));
    s->adam_t = 0;
    return MSD_OK;
}

extern "C" int msd_surrogate_train_step(msd_surrogate* s, double lr, double* loss)
{
    NONNULL(s);
    if (int rc = surrogate_loss_grads(s, loss)) return rc;
    return msd_surrogate_adam(s, lr);
}

// max |a - b| and max |b| over [3][n] device fields
int field_maxdiff(const float* a, const float* b, size_t n3, double* maxdiff, double* maxb)
{
    const int blocks = 512;
    if (!G.part)    // kept for the process: the check runs at every learned window
        if (int rc = dalloc(&G.part, blocks)) return rc;
    hipError_t e = launch_cnn_maxdiff(n3, a, b, G.part, blocks, G.stream);
    std::vector<float2> h(blocks);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), G.part, sizeof(float2) * blocks, hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "field check: %s", hipGetErrorString(e));
    double md = 0.0, mb = 0.0;
    for (const float2& v : h) { md = std::max(md, (double)v.x); mb = std::max(mb, (double)v.y); }
    *maxdiff = md;
    *maxb = mb;
    return MSD_OK;
}

// Surrogate prediction of the dipole field at m (m_ref, H_ref its reference
// state and exact field) against a field given on the host: returns
// max|pred - exact| / max|exact| (the residual the step driver checks).
static int surrogate_field_error(msd_surrogate* s, ExactField f, const msd_system* m, const msd_system* m_ref,
                                 double* rel_err, double* frozen_err)
{
    if (!same_dims(s->d, m->d.nx, m->d.ny, m->d.nz) || !same_dims(s->d, m_ref->d.nx, m_ref->d.ny, m_ref->d.nz) ||
        !f.fits(s->d))
        return fail(MSD_EMISMATCH, "Surrogate.fieldError: SpinSystem or field operator size differs");
    const size_t n3 = (size_t)3 * s->d.n;
    // exact H(m) into T, exact H(m_ref) into D, prediction D + CNN(m - m_ref) into Y3
    if (int rc = f.of(m->S, s->T)) return rc;
    if (int rc = f.of(m_ref->S, s->D)) return rc;
    hipError_t e = launch_cnn_input(s->d, m->S, m_ref->S, (float)s->s_in, s->X0, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.fieldError: %s", hipGetErrorString(e));
    if (int rc = s->forward(s->Y3, s->D, G.stream)) return rc;
    double md, mb;
    if (int rc = field_maxdiff(s->Y3, s->T, n3, &md, &mb)) return rc;
    *rel_err = mb > 0 ? md / mb : md;
    if (frozen_err) {
        if (int rc = field_maxdiff(s->D, s->T, n3, &md, &mb)) return rc;
        *frozen_err = mb > 0 ? md / mb : md;
    }
    return MSD_OK;
}

extern "C" int msd_surrogate_field_error(msd_surrogate* s, msd_dipole* dip, const msd_system* m, const msd_system* m_ref,
                                         double* rel_err, double* frozen_err)
{
    NONNULL(s); NONNULL(dip); NONNULL(m); NONNULL(m_ref); NONNULL(rel_err);
    return surrogate_field_error(s, ExactField{dip, nullptr}, m, m_ref, rel_err, frozen_err);
}

extern "C" int msd_surrogate_field_error_longrange3d(msd_surrogate* s, msd_longrange3d* lr, const msd_system* m,
                                                     const msd_system* m_ref, double* rel_err, double* frozen_err)
{
    NONNULL(s); NONNULL(lr); NONNULL(m); NONNULL(m_ref); NONNULL(rel_err);
    return surrogate_field_error(s, ExactField{nullptr, lr}, m, m_ref, rel_err, frozen_err);
}
