// ewpt_brick.hip -- fused RK4 stage kernel, LDS-tiled.
//
// One workgroup owns a TX*TY*TZ brick of sites, one site per thread.  The 32
// stencilled channels of the stage input are streamed one at a time through a
// double-buffered LDS tile holding the brick plus the star-shaped halo the
// radius-3 stencils of diff_scheme.jl:157-311 touch (the six 3-deep face
// slabs; edges and corners are never read).  The stage input y + c*k is
// formed when a value is staged (ewpt_stage.h), so the reference's broadcast
// temporaries (main_rk4.jl:6226-6242 ...) never exist.  Periodic wrap is done
// on the global index; there is no halo exchange.
//
// The right-hand side of rk4_step! (main_rk4.jl:337-657) is linear in the
// spatial derivatives once the 36 centre values are known.  Each channel's
// derivatives are therefore folded into 20 output accumulators (plus the
// magnetic-field accumulators B^a, B^Y and u_j = phi.d_j phi) as soon as they
// are computed, instead of holding all 80 derivative values.  The formulas
// are those of ewpt_point.h, distributed over the derivative terms.
//
// Stage S=4 also finishes the step: y' = k4*(dt/6) + n with
// n = ((k1*(dt/6) + y) + k2*(dt/3)) + k3*(dt/3) (main_rk4.jl:6400-6437).
#include <type_traits>
#include <utility>

#include "ewpt_internal.h"

namespace ewpt {

namespace {

template <int TX, int TY, int TZ>
struct BrickShape {
    static constexpr int NT = TX * TY * TZ;
    static constexpr int HX = TX + 6, HY = TY + 6, HZ = TZ + 6;
    static constexpr int BOX = HX * HY * HZ;   // LDS doubles per buffer
    static constexpr int NXF = 6 * TY * TZ;    // x-face halo elements
    static constexpr int NYF = 6 * TX * TZ;
    static constexpr int NZF = 6 * TX * TY;
};

struct BrickGeo {
    int nx, ny, nz;
    int nbx, nby, nbz;      // bricks per axis
    int xcd_map;            // 1: remap block ids so each XCD walks a compact region
    double inv60dx, invdx2;
};

// stencilled channel of pass p (0..31): state channel and which halo it needs
// (0 = full star, 1/2/3 = x/y/z faces only, i.e. one-direction derivative)
__host__ __device__ constexpr int pass_channel(int p)
{
    return p < 4 ? CH_PHI + p                       // phi_1..4
         : p < 7 ? CH_Y + (p - 4)                   // Y_x, Y_y, Y_z
         : p < 16 ? CH_W + 3 * ((p - 7) % 3) + (p - 7) / 3   // W^a_i, component-major
         : p < 20 ? CH_GAM + (p - 16)               // Gamma^1..3, Sigma
         : p < 32 ? ((p - 20) % 4 < 3 ? CH_E + 3 * ((p - 20) % 4) + (p - 20) / 4 : CH_EY + (p - 20) / 4)
         : -1;
}
__host__ __device__ constexpr int pass_dir(int p) { return p < 20 ? 0 : 1 + (p - 20) / 4; }

__device__ __forceinline__ int wrapc(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// box coordinates of halo element f of the faces selected by dir
template <int TX, int TY, int TZ>
__device__ __forceinline__ void face_elem(int dir, int f, int &lx, int &ly, int &lz)
{
    using B = BrickShape<TX, TY, TZ>;
    int d = dir;
    if (dir == 0) {
        if (f < B::NXF) d = 1;
        else if (f < B::NXF + B::NYF) { d = 2; f -= B::NXF; }
        else { d = 3; f -= B::NXF + B::NYF; }
    }
    if (d == 1) {
        const int h = f % 6, r = f / 6;
        lx = h < 3 ? h : TX + h;
        ly = 3 + r % TY;
        lz = 3 + r / TY;
    } else if (d == 2) {
        const int h = (f / TX) % 6;
        lx = 3 + f % TX;
        ly = h < 3 ? h : TY + h;
        lz = 3 + f / (TX * 6);
    } else {
        const int h = f / (TX * TY);
        lx = 3 + f % TX;
        ly = 3 + (f / TX) % TY;
        lz = h < 3 ? h : TZ + h;
    }
}

template <int TX, int TY, int TZ>
__host__ __device__ constexpr int face_count(int dir)
{
    using B = BrickShape<TX, TY, TZ>;
    return dir == 0 ? B::NXF + B::NYF + B::NZF : dir == 1 ? B::NXF : dir == 2 ? B::NYF : B::NZF;
}

// column b of A_i = h(W1 T1 + W2 T2 + W3 T3) + hy Y T_Y applied to val, added
// to acc (rows read off apply_gen in ewpt_point.h)
template <int b>
__device__ __forceinline__ void add_gen_col(double h, double W1, double W2, double W3, double hy, double Yv,
                                            double val, double acc[4])
{
    if constexpr (b == 0) {
        acc[1] -= (h * W3 + hy * Yv) * val; acc[2] += h * W2 * val; acc[3] -= h * W1 * val;
    } else if constexpr (b == 1) {
        acc[0] += (h * W3 + hy * Yv) * val; acc[2] += h * W1 * val; acc[3] += h * W2 * val;
    } else if constexpr (b == 2) {
        acc[0] -= h * W2 * val; acc[1] -= h * W1 * val; acc[3] += (h * W3 - hy * Yv) * val;
    } else {
        acc[0] += h * W1 * val; acc[1] -= h * W2 * val; acc[2] += (hy * Yv - h * W3) * val;
    }
}

// coefficient of v_b in phi.T_a v (a = 0,1,2) and phi.T_Y v
template <int a, int b>
__device__ __forceinline__ double phiT_coef(const double p[4])
{
    constexpr int tab[3][4][2] = {{{3, -1}, {2, 1}, {1, -1}, {0, 1}},
                                  {{2, 1}, {3, 1}, {0, -1}, {1, -1}},
                                  {{1, -1}, {0, 1}, {3, 1}, {2, -1}}};
    return tab[a][b][1] > 0 ? p[tab[a][b][0]] : -p[tab[a][b][0]];
}
template <int b>
__device__ __forceinline__ double phiTY_coef(const double p[4])
{
    constexpr int tab[4][2] = {{1, -1}, {0, 1}, {3, -1}, {2, 1}};
    return tab[b][1] > 0 ? p[tab[b][0]] : -p[tab[b][0]];
}

// Load/store at a 32-bit byte offset from a (uniform) channel base pointer:
// lets the compiler use the SGPR-base + VGPR-offset addressing form instead
// of materialising one 64-bit address per channel.
template <typename T>
__device__ __forceinline__ T ldo(const T *base, unsigned boff)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void sto(T *base, unsigned boff, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff) = v;
}

// Opaque use/def of a register value: keeps the compiler from sinking the
// arithmetic that produced it past the next barrier (which would keep every
// staged LDS value alive across passes).
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }

// sign of eps^{a b c} for the third index c
__host__ __device__ constexpr int eps3(int a, int b) { return ((b - a + 3) % 3 == 1) ? 1 : -1; }

template <int TX, int TY, int TZ, int S>
__global__ void __launch_bounds__(TX *TY *TZ) stage_brick(StageBufs<double> b, BrickGeo g, Phys P, double dt)
{
    using B = BrickShape<TX, TY, TZ>;
    // dynamic LDS: two staging buffers, then (stage 4) the 20 partial sums n
    extern __shared__ double lds_dyn[];
    double *lds[2] = {lds_dyn, lds_dyn + B::BOX};
    double *nstash = lds_dyn + 2 * B::BOX;

    // ---- brick of this workgroup
    int bid = blockIdx.x;
    const int nbrick = g.nbx * g.nby * g.nbz;
    if (g.xcd_map) {
        // consecutive ids go to different XCDs; give XCD x the contiguous range
        // [x*nbrick/8, (x+1)*nbrick/8) so its resident workgroups are neighbours
        const int per = nbrick >> 3;
        bid = (bid & 7) * per + (bid >> 3);
    }
    const int bx = bid % g.nbx, by = (bid / g.nbx) % g.nby, bz = bid / (g.nbx * g.nby);
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
    const int t = threadIdx.x;
    const int tx = t % TX, ty = (t / TX) % TY, tz = t / (TX * TY);
    const long nxl = g.nx, nxy = (long)g.nx * g.ny;
    const unsigned qb = (unsigned)((x0 + tx) + nxl * (y0 + ty) + nxy * (z0 + tz)) * 8u;
    auto Y = [&](int ch) { return b.y + (long)ch * b.n3; };
    auto K = [&](const double *k, int m) { return k + (long)m * b.n3; };

    // ---- stage 4: write the field part of y' now, park n for channels 16..35
    // in LDS until k4 is known
    if constexpr (S == 4) {
        const double d6 = dt / 6.0, d3 = dt / 3.0;
        // groups of four channels, so the scheduler does not put all ~160
        // loads of this block in flight at once
#pragma unroll
        for (int ch = 0; ch < 16; ch++) {
            const double yv = ldo(Y(ch), qb);
            const double piv = ldo(Y(ch + 20), qb);
            const double k2 = piv + ldo(K(b.k1, ch + 4), qb) * (dt * 0.5);
            const double k3 = piv + ldo(K(b.k2, ch + 4), qb) * (dt * 0.5);
            const double k4 = piv + ldo(K(b.k3, ch + 4), qb) * dt;
            double n = piv * d6 + yv;
            n = k2 * d3 + n;
            n = k3 * d3 + n;
            sto(b.ynew + (long)ch * b.n3, qb, k4 * d6 + n);
            if (ch % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int m = 0; m < NK; m++) {
            double n = ldo(K(b.k1, m), qb) * d6 + ldo(Y(16 + m), qb);
            n = ldo(K(b.k2, m), qb) * d3 + n;
            nstash[m * B::NT + t] = ldo(K(b.k3, m), qb) * d3 + n;
            if (m % 4 == 3) __builtin_amdgcn_sched_barrier(0);
        }
    }

    __builtin_amdgcn_sched_barrier(0);

    // ---- centre values of the stage input (ewpt_stage.h)
    double c[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
        const double yv = ldo(Y(ch), qb);
        if constexpr (S == 1) {
            c[ch] = yv;
        } else if (ch < 16) {
            const double piv = ldo(Y(ch + 20), qb);
            if constexpr (S == 2) c[ch] = yv + piv * (dt * 0.5);
            if constexpr (S == 3) c[ch] = yv + (piv + ldo(K(b.k1, ch + 4), qb) * (dt * 0.5)) * (dt * 0.5);
            if constexpr (S == 4) c[ch] = yv + (piv + ldo(K(b.k2, ch + 4), qb) * (dt * 0.5)) * dt;
        } else {
            const double *kp = (S == 2) ? b.k1 : (S == 3) ? b.k2 : b.k3;
            c[ch] = yv + ldo(K(kp, ch - 16), qb) * (S == 4 ? dt : dt * 0.5);
        }
    }

    const double gw = P.gw, gy = P.gy, gp2 = P.gp2, h = 0.5 * gw, hy = 0.5 * gy;
    const double p[4] = {c[0], c[1], c[2], c[3]};
    auto W = [&](int a, int i) -> double { return c[CH_W + 3 * a + i]; };
    auto E = [&](int a, int i) -> double { return c[CH_E + 3 * a + i]; };
    const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    const double phipi = c[CH_PI] * p[0] + c[CH_PI + 1] * p[1] + c[CH_PI + 2] * p[2] + c[CH_PI + 3] * p[3];

    // ---- accumulators
    double ktphi[4], ktW[3][3], ktY[3], kG[3], kS, u[3] = {0.0, 0.0, 0.0};
    double Bw[3][3], BYv[3] = {0.0, 0.0, 0.0};
    {
        const double pot = 2.0 * P.lam * (r2 - P.vev * P.vev) + 0.5 * P.jex;
        double Gp[4];
        apply_gen(h, c[CH_GAM], c[CH_GAM + 1], c[CH_GAM + 2], hy, c[CH_SIG], p, Gp);
        const double inv_r2 = 1.0 / r2;   // keeps the reference's 0/0 at |phi| = 0
#pragma unroll
        for (int bb = 0; bb < 4; bb++)
            ktphi[bb] = Gp[bb] - pot * p[bb] - P.gam * p[bb] * (p[bb] * c[CH_PI + bb]) * inv_r2;
        const double pi[4] = {c[CH_PI], c[CH_PI + 1], c[CH_PI + 2], c[CH_PI + 3]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double na = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) na -= xprod(a, W(0, i), W(1, i), W(2, i), E(0, i), E(1, i), E(2, i));
            kG[a] = gp2 * gw * (na + phiT(a, p, pi));
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ktW[a][i] = -gw * xprod(a, W(0, i), W(1, i), W(2, i), c[CH_GAM], c[CH_GAM + 1], c[CH_GAM + 2]);
                const int j = (i + 1) % 3, k = (i + 2) % 3;
                Bw[a][i] = gw * xprod(a, W(0, j), W(1, j), W(2, j), W(0, k), W(1, k), W(2, k));
            }
        }
        kS = gp2 * gy * phiTY(p, pi);
#pragma unroll
        for (int i = 0; i < 3; i++) ktY[i] = 0.0;
    }

    // ---- staging of one channel into LDS: centre from registers, faces from
    // global.  Per face class (0 = star, 1/2/3 = x/y/z faces) each thread owns
    // up to MAXF halo elements; their in-channel global offsets and LDS
    // offsets are computed once (-1 marks an empty slot).
    constexpr int MAXF = (face_count<TX, TY, TZ>(0) + B::NT - 1) / B::NT;
    const double inv60dx = g.inv60dx, invdx2 = g.invdx2;
    int goff[4][MAXF], loff[4][MAXF];   // goff: byte offset within a channel, -1 = empty
#pragma unroll
    for (int dir = 0; dir < 4; dir++) {
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            const int f = t + s * B::NT;
            goff[dir][s] = -1;
            loff[dir][s] = 0;
            if (f < face_count<TX, TY, TZ>(dir)) {
                int lx, ly, lz;
                face_elem<TX, TY, TZ>(dir, f, lx, ly, lz);
                goff[dir][s] = (int)((wrapc(x0 + lx - 3, g.nx) + nxl * wrapc(y0 + ly - 3, g.ny) +
                                      nxy * wrapc(z0 + lz - 3, g.nz)) * 8);
                loff[dir][s] = lz * B::HX * B::HY + ly * B::HX + lx;
            }
        }
    }
    const int lcen = (tz + 3) * B::HX * B::HY + (ty + 3) * B::HX + (tx + 3);

    auto fetch = [&](auto pc, double (&v)[MAXF][3]) {
        constexpr int pp = decltype(pc)::value;
        constexpr int ch = pass_channel(pp);
        constexpr int dir = pass_dir(pp);
        constexpr int nf = face_count<TX, TY, TZ>(dir);
        const double *y0p = b.y + (long)ch * b.n3;
        const double *y1p = nullptr, *k2p = nullptr;
        if constexpr (S > 1) {
            if constexpr (ch < 16) {
                y1p = b.y + (long)(ch + 20) * b.n3;
                if constexpr (S == 3) k2p = b.k1 + (long)(ch + 4) * b.n3;
                if constexpr (S == 4) k2p = b.k2 + (long)(ch + 4) * b.n3;
            } else {
                const double *kp = (S == 2) ? b.k1 : (S == 3) ? b.k2 : b.k3;
                y1p = kp + (long)(ch - 16) * b.n3;
            }
        }
#pragma unroll
        for (int s = 0; s < MAXF; s++) {
            if (s * B::NT < nf) {
                const int o = goff[dir][s];
                if (o >= 0) {
                    v[s][0] = ldo(y0p, (unsigned)o);
                    if constexpr (S > 1) v[s][1] = ldo(y1p, (unsigned)o);
                    if constexpr (S > 2 && ch < 16) v[s][2] = ldo(k2p, (unsigned)o);
                }
            }
        }
    };
    auto stage_value = [&](auto pc, const double (&v)[3]) -> double {
        constexpr int ch = pass_channel(decltype(pc)::value);
        if constexpr (S == 1) return v[0];
        else if constexpr (ch < 16) {
            if constexpr (S == 2) return v[0] + v[1] * (dt * 0.5);
            else if constexpr (S == 3) return v[0] + (v[1] + v[2] * (dt * 0.5)) * (dt * 0.5);
            else return v[0] + (v[1] + v[2] * (dt * 0.5)) * dt;
        } else {
            return v[0] + v[1] * (S == 4 ? dt : dt * 0.5);
        }
    };
    auto store = [&](auto pc, const double (&v)[MAXF][3], double *buf) {
        constexpr int pp = decltype(pc)::value;
        constexpr int dir = pass_dir(pp);
        constexpr int nf = face_count<TX, TY, TZ>(dir);
        buf[lcen] = c[pass_channel(pp)];
#pragma unroll
        for (int s = 0; s < MAXF; s++)
            if (s * B::NT < nf && goff[dir][s] >= 0) buf[loff[dir][s]] = stage_value(pc, v[s]);
    };

    // derivatives of the staged channel at this thread's site
    auto deriv = [&](const double *buf, int ax, double cen, bool second, double &d1, double &d2) {
        const int stride = ax == 0 ? 1 : ax == 1 ? B::HX : B::HX * B::HY;
        const double *o = buf + lcen;
        const double m3 = o[-3 * stride], m2 = o[-2 * stride], m1 = o[-stride];
        const double a1 = o[stride], a2 = o[2 * stride], a3 = o[3 * stride];
        d1 = sten_d1(m3, m2, m1, a1, a2, a3, inv60dx);
        if (second) d2 = sten_d2(m3, m2, m1, cen, a1, a2, a3, invdx2);
    };

    // fold the derivatives of pass pp into the accumulators
    auto consume = [&](auto pc, const double *buf) {
        constexpr int pp = decltype(pc)::value;
        constexpr int ch = pass_channel(pp);
        const double cen = c[ch];
        if constexpr (pp < 4) {
            // phi_b: kt_phi (Laplacian, A_i(d_i phi + D_i phi)), phi.T D into kt_W, kt_Y, u_j
            constexpr int bb = pp;
            double lap = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double d1, d2;
                deriv(buf, i, cen, true, d1, d2);
                lap += d2;
                double Ap[4];
                apply_gen(h, W(0, i), W(1, i), W(2, i), hy, c[CH_Y + i], p, Ap);
                const double D = d1 + Ap[bb];
                add_gen_col<bb>(h, W(0, i), W(1, i), W(2, i), hy, c[CH_Y + i], d1 + D, ktphi);
                ktW[0][i] += gw * phiT_coef<0, bb>(p) * D;
                ktW[1][i] += gw * phiT_coef<1, bb>(p) * D;
                ktW[2][i] += gw * phiT_coef<2, bb>(p) * D;
                ktY[i] += gy * phiTY_coef<bb>(p) * D;
                u[i] += p[bb] * d1;
            }
            ktphi[bb] += lap;
        } else if constexpr (pp < 7) {
            // Y_i: Laplacian into kt_Y, d_j Y_i into B^Y
            constexpr int i = pp - 4;
            double lap = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double d1, d2;
                deriv(buf, j, cen, true, d1, d2);
                lap += d2;
                if (j != i) {
                    if (i == (j + 1) % 3) BYv[(j + 2) % 3] += d1;
                    else BYv[(j + 1) % 3] -= d1;
                }
            }
            ktY[i] += lap;
            if constexpr (pp == 6) {
                // B^Y complete: beta_Y terms (main_rk4.jl:354, 553-581)
                const double QY = c[CH_EY] * BYv[0] + c[CH_EY + 1] * BYv[1] + c[CH_EY + 2] * BYv[2];
#pragma unroll
                for (int bq = 0; bq < 4; bq++) ktphi[bq] += P.by * QY * p[bq];
#pragma unroll
                for (int ii = 0; ii < 3; ii++) {
                    const int j = (ii + 1) % 3, k = (ii + 2) % 3;
                    ktY[ii] += 2.0 * P.by * (u[j] * c[CH_EY + k] - u[k] * c[CH_EY + j] - phipi * BYv[ii]);
                }
            }
        } else if constexpr (pp < 16) {
            // W^a_i: Laplacian, -gw eps W^c_j d_j W^a_i into kt_W^{a'}_i, d_j W^a_i into B^a
            constexpr int i = (pp - 7) / 3, a = (pp - 7) % 3;
            double lap = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double d1, d2;
                deriv(buf, j, cen, true, d1, d2);
                lap += d2;
                // (d_j W_i x W_j)^{a'} = eps^{a' a c} d_j W^a_i W^c_j
                constexpr int a1 = (a + 1) % 3, a2 = (a + 2) % 3;   // a' values
                ktW[a1][i] -= gw * (double)eps3(a1, a) * W(3 - a1 - a, j) * d1;
                ktW[a2][i] -= gw * (double)eps3(a2, a) * W(3 - a2 - a, j) * d1;
                if (j != i) {
                    if (i == (j + 1) % 3) Bw[a][(j + 2) % 3] += d1;
                    else Bw[a][(j + 1) % 3] -= d1;
                }
            }
            ktW[a][i] += lap;
            if constexpr (pp == 15) {
                // B^a complete: -gw sum_j (W_j x F_ij)^a, beta_W terms
                double QW = 0.0;
#pragma unroll
                for (int aa = 0; aa < 3; aa++) QW += E(aa, 0) * Bw[aa][0] + E(0, 1) * Bw[aa][1] + E(0, 2) * Bw[aa][2];
#pragma unroll
                for (int bq = 0; bq < 4; bq++) ktphi[bq] += P.bw * QW * p[bq];
#pragma unroll
                for (int aa = 0; aa < 3; aa++) {
#pragma unroll
                    for (int ii = 0; ii < 3; ii++) {
                        const int j1 = (ii + 1) % 3, j2 = (ii + 2) % 3;
                        // F_{i,i+1} = B_{i+2}, F_{i,i+2} = -B_{i+1}
                        const double t2 = xprod(aa, W(0, j1), W(1, j1), W(2, j1), Bw[0][j2], Bw[1][j2], Bw[2][j2]) -
                                          xprod(aa, W(0, j2), W(1, j2), W(2, j2), Bw[0][j1], Bw[1][j1], Bw[2][j1]);
                        ktW[aa][ii] += -gw * t2 +
                                       2.0 * P.bw * (u[j1] * E(aa, j2) - u[j2] * E(aa, j1) - phipi * Bw[aa][ii]);
                    }
                }
            }
        } else if constexpr (pp < 20) {
            // Gamma^a, Sigma: gradients
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double d1, d2;
                deriv(buf, i, cen, false, d1, d2);
                if constexpr (pp < 19) ktW[pp - 16][i] -= d1;
                else ktY[i] -= d1;
            }
        } else {
            // E^a_i, E^Y_i: d_i only, divergences into k_Gamma, k_Sigma
            constexpr int i = (pp - 20) / 4, a = (pp - 20) % 4;
            double d1, d2;
            deriv(buf, i, cen, false, d1, d2);
            if constexpr (a < 3) kG[a] += (1.0 - gp2) * d1;
            else kS += (1.0 - gp2) * d1;
        }
    };

    // pin the accumulators still live after pass pp
    auto pin_all = [&](auto pc) {
        constexpr int pp = decltype(pc)::value;
#pragma unroll
        for (int i = 0; i < 4; i++) pin(ktphi[i]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            pin(ktY[i]); pin(kG[i]);
            if constexpr (pp < 15) pin(u[i]);
            if constexpr (pp < 6) pin(BYv[i]);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                pin(ktW[i][j]);
                if constexpr (pp < 15) pin(Bw[i][j]);
            }
        }
        pin(kS);
    };

    // ---- the channel pipeline: fetch pass p+1 while computing pass p
    double v[MAXF][3];
    fetch(std::integral_constant<int, 0>{}, v);
    store(std::integral_constant<int, 0>{}, v, lds[0]);
    __syncthreads();
    auto body = [&](auto pc) {
        constexpr int pp = decltype(pc)::value;
        if constexpr (pp + 1 < 32) fetch(std::integral_constant<int, pp + 1>{}, v);
        consume(pc, lds[pp & 1]);
        pin_all(pc);
        if constexpr (pp + 1 < 32) store(std::integral_constant<int, pp + 1>{}, v, lds[(pp + 1) & 1]);
        __syncthreads();
    };
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (body(std::integral_constant<int, I>{}), ...);
    }(std::make_integer_sequence<int, 32>{});

    // ---- outputs: k_Gamma(3), k_Sigma, kt_phi(4), kt_W(9), kt_Y(3)
    double o[NK];
#pragma unroll
    for (int a = 0; a < 3; a++) o[a] = kG[a];
    o[3] = kS;
#pragma unroll
    for (int bb = 0; bb < 4; bb++) o[4 + bb] = ktphi[bb];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) o[8 + 3 * a + i] = ktW[a][i];
#pragma unroll
    for (int i = 0; i < 3; i++) o[17 + i] = ktY[i];

    if constexpr (S < 4) {
#pragma unroll
        for (int m = 0; m < NK; m++) sto(b.kout + (long)m * b.n3, qb, o[m]);
    } else {
        const double d6 = dt / 6.0;
#pragma unroll
        for (int m = 0; m < NK; m++) sto(b.ynew + (long)(16 + m) * b.n3, qb, o[m] * d6 + nstash[m * B::NT + t]);
    }
}

template <int TX, int TY, int TZ>
bool brick_fits(const ewpt_lattice &L)
{
    // byte offsets within one channel are 32-bit
    return L.nx % TX == 0 && L.ny % TY == 0 && L.nz % TZ == 0 && L.nx >= TX && L.ny >= TY && L.nz >= TZ &&
           (long)L.nx * L.ny * L.nz * 8 < (1L << 31);
}

template <int TX, int TY, int TZ>
hipError_t launch_tile(int stage, const StageBufs<double> &b, const ewpt_lattice &L, const Phys &P, double dt,
                       int xcd_map, hipStream_t st)
{
    BrickGeo g;
    g.nx = L.nx; g.ny = L.ny; g.nz = L.nz;
    g.nbx = L.nx / TX; g.nby = L.ny / TY; g.nbz = L.nz / TZ;
    const int nb = g.nbx * g.nby * g.nbz;
    g.xcd_map = (xcd_map && nb % 8 == 0) ? 1 : 0;
    g.inv60dx = 1.0 / (60.0 * L.dx);
    g.invdx2 = 1.0 / (L.dx * L.dx);
    using B = BrickShape<TX, TY, TZ>;
    const dim3 grid(nb), block(TX * TY * TZ);
    const size_t lds13 = 2 * B::BOX * sizeof(double);
    const size_t lds4 = lds13 + NK * B::NT * sizeof(double);
    static bool attr_set = false;   // the 4-stage kernels need more than 64 KB of LDS
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&stage_brick<TX, TY, TZ, 4>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds4);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    switch (stage) {
    case 1: stage_brick<TX, TY, TZ, 1><<<grid, block, lds13, st>>>(b, g, P, dt); break;
    case 2: stage_brick<TX, TY, TZ, 2><<<grid, block, lds13, st>>>(b, g, P, dt); break;
    case 3: stage_brick<TX, TY, TZ, 3><<<grid, block, lds13, st>>>(b, g, P, dt); break;
    default: stage_brick<TX, TY, TZ, 4><<<grid, block, lds4, st>>>(b, g, P, dt); break;
    }
    return hipGetLastError();
}

}  // namespace

// Brick variants: 1 = 16x8x4, 2 = 8x8x8, 3 = 32x4x4, 4 = 16x16x2 (all with
// the XCD-aware block order); 5 = 16x8x4 in plain block order.
int brick_variant_count() { return 5; }

bool brick_supported(int variant, const ewpt_lattice &L)
{
    switch (variant) {
    case 1: case 5: return brick_fits<16, 8, 4>(L);
    case 2: return brick_fits<8, 8, 8>(L);
    case 3: return brick_fits<32, 4, 4>(L);
    case 4: return brick_fits<16, 16, 2>(L);
    default: return false;
    }
}

hipError_t launch_stage_brick(int variant, int stage, const StageBufs<double> &b, const ewpt_lattice &L,
                              const Phys &P, double dt, hipStream_t st)
{
    switch (variant) {
    case 1: return launch_tile<16, 8, 4>(stage, b, L, P, dt, 1, st);
    case 2: return launch_tile<8, 8, 8>(stage, b, L, P, dt, 1, st);
    case 3: return launch_tile<32, 4, 4>(stage, b, L, P, dt, 1, st);
    case 4: return launch_tile<16, 16, 2>(stage, b, L, P, dt, 1, st);
    case 5: return launch_tile<16, 8, 4>(stage, b, L, P, dt, 0, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ewpt
