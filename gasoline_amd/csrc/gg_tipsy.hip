// Tipsy files in and out of the store: the header (tipsydefs.h:38-45 read
// and written by msrReadTipsy/msrWriteTipsy), pkdReadTipsy (pkd.c:297-640)
// and pkdWriteTipsy (pkd.c:1628-1800) for one process holding every
// particle; and the ASCII output arrays of msrOutArray/msrOutVector
// (master.c:5420-5545, pkdOutVector outtype.c:806-1000, iBinaryOutput 0).
// Standard files are XDR (big-endian; a 32-byte header with a pad word; gas
// 48, dark 36, star 44 bytes per particle), native ones the host's structs
// (sizeof(struct dump) = 32 on x86-64).  Host code: the file is read and
// written in chunks; the fields move with gg_put / gg_get.
#include "gg_internal.h"

#include <cfloat>
#include <cstdio>
#include <cstring>
#include <vector>

namespace gg {
namespace {

constexpr int GAS_W = 12, DARK_W = 9, STAR_W = 11;   // floats per particle record

struct TipsyHeader {
    double time;
    int nbodies, ndim, nsph, ndark, nstar;
};

uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

// Host-order floats from the file's byte order (in place).
void decode_floats(float *f, size_t n, bool xdr)
{
    if (!xdr) return;
    uint32_t *u = reinterpret_cast<uint32_t *>(f);
    for (size_t k = 0; k < n; ++k) u[k] = bswap32(u[k]);
}

int read_header(FILE *fp, bool xdr, TipsyHeader *h)
{
    unsigned char b[32];
    if (fread(b, 1, 32, fp) != 32) return GG_EINVAL;
    if (xdr) {
        uint64_t t = 0;
        for (int k = 0; k < 8; ++k) t = (t << 8) | b[k];
        std::memcpy(&h->time, &t, 8);
        int *dst[5] = {&h->nbodies, &h->ndim, &h->nsph, &h->ndark, &h->nstar};
        for (int k = 0; k < 5; ++k) {
            uint32_t v;
            std::memcpy(&v, b + 8 + 4 * k, 4);
            v = bswap32(v);
            std::memcpy(dst[k], &v, 4);
        }
    } else {
        std::memcpy(&h->time, b, 8);
        std::memcpy(&h->nbodies, b + 8, 4);
        std::memcpy(&h->ndim, b + 12, 4);
        std::memcpy(&h->nsph, b + 16, 4);
        std::memcpy(&h->ndark, b + 20, 4);
        std::memcpy(&h->nstar, b + 24, 4);
    }
    return GG_OK;
}

void encode_header(const TipsyHeader &h, bool xdr, unsigned char *b)
{
    std::memset(b, 0, 32);
    if (xdr) {
        uint64_t t;
        std::memcpy(&t, &h.time, 8);
        for (int k = 7; k >= 0; --k) {
            b[k] = (unsigned char)(t & 0xff);
            t >>= 8;
        }
        const int src[5] = {h.nbodies, h.ndim, h.nsph, h.ndark, h.nstar};
        for (int k = 0; k < 5; ++k) {
            const uint32_t v = bswap32((uint32_t)src[k]);
            std::memcpy(b + 8 + 4 * k, &v, 4);
        }
    } else {
        std::memcpy(b, &h.time, 8);
        std::memcpy(b + 8, &h.nbodies, 4);
        std::memcpy(b + 12, &h.ndim, 4);
        std::memcpy(b + 16, &h.nsph, 4);
        std::memcpy(b + 20, &h.ndark, 4);
        std::memcpy(b + 24, &h.nstar, 4);
    }
}

struct HostStore {   // the fields a Tipsy file carries, caller order
    std::vector<double> r, v, vPred, m, soft, rho, u, uPred, pot, dt, dtNew, fb2;
    std::vector<uint32_t> bits;
    std::vector<int> rung;
    void resize(int n)
    {
        r.assign(3 * (size_t)n, 0.0);
        v.assign(3 * (size_t)n, 0.0);
        vPred.assign(3 * (size_t)n, 0.0);
        for (auto *a : {&m, &soft, &rho, &u, &uPred, &pot, &fb2}) a->assign(n, 0.0);
        dt.assign(n, (double)FLT_MAX);
        dtNew.assign(n, (double)FLT_MAX);
        bits.assign(n, 0);
        rung.assign(n, 0);
    }
};

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_tipsy_header(const char *path, int bStandard, double *dTime, int *nBodies,
                               int *nGas, int *nDark, int *nStar)
{
    GG_CHECK(path, GG_EINVAL, "null path");
    FILE *fp = fopen(path, "rb");
    GG_CHECK(fp, GG_EINVAL, "cannot open %s", path);
    TipsyHeader h;
    const int rc = read_header(fp, bStandard != 0, &h);
    fclose(fp);
    GG_CHECK(rc == GG_OK, GG_EINVAL, "%s: short header", path);
    if (dTime) *dTime = h.time;
    if (nBodies) *nBodies = h.nbodies;
    if (nGas) *nGas = h.nsph;
    if (nDark) *nDark = h.ndark;
    if (nStar) *nStar = h.nstar;
    return GG_OK;
}

extern "C" int gg_read_tipsy(GGPKD pkd, const char *path, int bStandard, double dvFac,
                             double dTuFac, double *pdTime)
{
    GG_CHECK(pkd && path, GG_EINVAL, "null argument");
    const bool xdr = bStandard != 0;
    FILE *fp = fopen(path, "rb");
    GG_CHECK(fp, GG_EINVAL, "cannot open %s", path);
    TipsyHeader h;
    if (read_header(fp, xdr, &h) != GG_OK || h.nsph < 0 || h.ndark < 0 || h.nstar < 0 ||
        h.nbodies != h.nsph + h.ndark + h.nstar || h.nbodies > pkd->nStore) {
        fclose(fp);
        set_error("%s: bad header, or %d particles for a store of %d", path, h.nbodies,
                  pkd->nStore);
        return GG_EINVAL;
    }
    const int n = h.nbodies;
    HostStore S;
    S.resize(n);
    // records in file order: gas, dark, star (iOrder = position, pkd.c:469)
    const int cnt[3] = {h.nsph, h.ndark, h.nstar}, width[3] = {GAS_W, DARK_W, STAR_W};
    const uint32_t type[3] = {GG_TYPE_GAS, GG_TYPE_DARK, GG_TYPE_STAR};
    std::vector<float> buf;
    int i0 = 0;
    int rc = GG_OK;
    for (int t = 0; t < 3 && rc == GG_OK; ++t) {
        const int w = width[t];
        const int chunk = 1 << 16;
        for (int a = 0; a < cnt[t] && rc == GG_OK; a += chunk) {
            const int m = std::min(chunk, cnt[t] - a);
            buf.resize((size_t)m * w);
            if (fread(buf.data(), sizeof(float), buf.size(), fp) != buf.size()) {
                set_error("%s: file ends inside the particle records", path);
                rc = GG_EINVAL;
                break;
            }
            decode_floats(buf.data(), buf.size(), xdr);
            for (int k = 0; k < m; ++k) {
                const float *f = buf.data() + (size_t)k * w;
                const size_t i = (size_t)i0 + a + k;
                S.m[i] = f[0];
                for (int j = 0; j < 3; ++j) {
                    S.r[3 * i + j] = f[1 + j];
                    S.v[3 * i + j] = dvFac * (double)f[4 + j];
                }
                S.bits[i] = type[t];
                if (t == 0) {   // gas: rho, temp, eps, metals, phi
                    for (int j = 0; j < 3; ++j) S.vPred[3 * i + j] = S.v[3 * i + j];
                    S.rho[i] = f[7];
                    S.u[i] = dTuFac * (double)f[8];
                    S.uPred[i] = S.u[i];
                    S.soft[i] = f[9];
                    S.pot[i] = f[11];
                } else if (t == 1) {   // dark: eps, phi
                    S.soft[i] = f[7];
                    S.pot[i] = f[8];
                } else {   // star: metals, tform, eps, phi
                    S.soft[i] = f[9];
                    S.pot[i] = f[10];
                }
            }
        }
        i0 += cnt[t];
    }
    fclose(fp);
    if (rc != GG_OK) return rc;
    GG_TRY(gg_set_local(pkd, n));
    if (pdTime) *pdTime = h.time;
    if (n == 0) return GG_OK;
    GG_TRY(gg_put(pkd, GG_R, S.r.data(), 0));
    GG_TRY(gg_put(pkd, GG_V, S.v.data(), 0));
    GG_TRY(gg_put(pkd, GG_VPRED, S.vPred.data(), 0));
    GG_TRY(gg_put(pkd, GG_MASS, S.m.data(), 0));
    GG_TRY(gg_put(pkd, GG_SOFT, S.soft.data(), 0));
    GG_TRY(gg_put(pkd, GG_DENSITY, S.rho.data(), 0));
    GG_TRY(gg_put(pkd, GG_U, S.u.data(), 0));
    GG_TRY(gg_put(pkd, GG_UPRED, S.uPred.data(), 0));
    GG_TRY(gg_put(pkd, GG_POT, S.pot.data(), 0));
    GG_TRY(gg_put(pkd, GG_FBALL2, S.fb2.data(), 0));
    GG_TRY(gg_put(pkd, GG_DT, S.dt.data(), 0));
    GG_TRY(gg_put(pkd, GG_DTNEW, S.dtNew.data(), 0));
    GG_TRY(gg_put(pkd, GG_IRUNG, S.rung.data(), 0));
    GG_TRY(gg_put(pkd, GG_IACTIVE, S.bits.data(), 0));
    if (pdTime) *pdTime = h.time;
    return GG_OK;
}

extern "C" int gg_write_tipsy(GGPKD pkd, const char *path, int bStandard, double dTime,
                              double dvFac, double duTFac)
{
    GG_CHECK(pkd && path, GG_EINVAL, "null argument");
    const bool xdr = bStandard != 0;
    const int n = pkd->nLocal;
    const size_t nn = (size_t)std::max(n, 1);   // non-empty buffers for gg_get
    std::vector<double> r(3 * nn), v(3 * nn), m(nn), soft(nn), rho(nn), u(nn), pot(nn);
    std::vector<uint32_t> bits(nn);
    GG_TRY(gg_get(pkd, GG_R, r.data(), 0));
    GG_TRY(gg_get(pkd, GG_V, v.data(), 0));
    GG_TRY(gg_get(pkd, GG_MASS, m.data(), 0));
    GG_TRY(gg_get(pkd, GG_SOFT, soft.data(), 0));
    GG_TRY(gg_get(pkd, GG_DENSITY, rho.data(), 0));
    GG_TRY(gg_get(pkd, GG_U, u.data(), 0));
    GG_TRY(gg_get(pkd, GG_POT, pot.data(), 0));
    GG_TRY(gg_get(pkd, GG_IACTIVE, bits.data(), 0));
    TipsyHeader h{dTime, n, 3, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const uint32_t b = bits[i];
        if (b & GG_TYPE_DARK) ++h.ndark;
        else if (b & GG_TYPE_GAS) ++h.nsph;
        else if (b & GG_TYPE_STAR) ++h.nstar;
        else GG_CHECK(false, GG_EINVAL, "particle %d has no type bit (gas, dark or star)", i);
    }
    FILE *fp = fopen(path, "wb");
    GG_CHECK(fp, GG_EINVAL, "cannot create %s", path);
    unsigned char hb[32];
    encode_header(h, xdr, hb);
    bool ok = fwrite(hb, 1, 32, fp) == 32;
    // gas, dark, star records in store order within each type (the
    // reference writes the store in iOrder order, which sorts them so)
    std::vector<float> rec;
    for (int t = 0; t < 3 && ok; ++t) {
        rec.clear();
        for (int i = 0; i < n; ++i) {
            const uint32_t b = bits[i];
            const int ti = (b & GG_TYPE_DARK) ? 1 : (b & GG_TYPE_GAS) ? 0 : 2;   // pkd.c:1655-1745
            if (ti != t) continue;
            float f[GAS_W];
            f[0] = (float)m[i];
            for (int j = 0; j < 3; ++j) {
                f[1 + j] = (float)r[3 * (size_t)i + j];
                const double vt = dvFac * v[3 * (size_t)i + j];
                f[4 + j] = (float)vt;
            }
            int w;
            if (t == 0) {   // rho, temp = duTFac u, eps, metals (not carried), phi
                const double vt = duTFac * u[i];
                f[7] = (float)rho[i];
                f[8] = (float)vt;
                f[9] = (float)soft[i];
                f[10] = 0.0f;
                f[11] = (float)pot[i];
                w = GAS_W;
            } else if (t == 1) {
                f[7] = (float)soft[i];
                f[8] = (float)pot[i];
                w = DARK_W;
            } else {   // metals and formation time are not carried
                f[7] = 0.0f;
                f[8] = 0.0f;
                f[9] = (float)soft[i];
                f[10] = (float)pot[i];
                w = STAR_W;
            }
            rec.insert(rec.end(), f, f + w);
            if (rec.size() >= (1u << 20)) {
                if (xdr) decode_floats(rec.data(), rec.size(), true);
                ok = ok && fwrite(rec.data(), sizeof(float), rec.size(), fp) == rec.size();
                rec.clear();
            }
        }
        if (xdr) decode_floats(rec.data(), rec.size(), true);
        ok = ok && fwrite(rec.data(), sizeof(float), rec.size(), fp) == rec.size();
    }
    ok = (fclose(fp) == 0) && ok;
    GG_CHECK(ok, GG_EINVAL, "writing %s failed", path);
    return GG_OK;
}

extern "C" int gg_out_array(GGPKD pkd, const char *path, int field)
{
    GG_CHECK(pkd && path, GG_EINVAL, "null argument");
    GG_CHECK(field >= 0 && field < GG_NFIELDS && !field_is32(field), GG_EINVAL,
             "field %d is not an FP64 field", field);
    const int n = pkd->nLocal, w = field_width(field);
    std::vector<double> v((size_t)w * std::max(n, 1));
    GG_TRY(gg_get(pkd, field, v.data(), 0));
    FILE *fp = fopen(path, "w");
    GG_CHECK(fp, GG_EINVAL, "cannot create %s", path);
    bool ok = fprintf(fp, "%d\n", n) > 0;
    // a vector is written a component at a time (iDim 0, 1, 2)
    for (int d = 0; d < w && ok; ++d)
        for (int i = 0; i < n && ok; ++i) ok = fprintf(fp, "%.14g\n", v[(size_t)w * i + d]) > 0;
    ok = (fclose(fp) == 0) && ok;
    GG_CHECK(ok, GG_EINVAL, "writing %s failed", path);
    return GG_OK;
}
