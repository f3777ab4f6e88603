// lbm_fused_pair.h -- fused D3Q19 step of the 16-bit storage forms with two
// z-adjacent nodes per thread.
//
// The march is that of k_fused_lp (lbm_fused_lp.h): a workgroup owns a
// TY x TZ tile of a y-z cross-section and walks a chunk of x planes, with
// psi and rho of planes x-1 .. x+2 (tile plus a one-node ring) in LDS and
// the pulls of plane x+2 in flight during the collision of plane x.  Here a
// thread owns the node pair (z, z+1), z even, so every gather and every
// store moves 4 bytes (two 16-bit values) instead of 2: a wavefront
// instruction then carries 256 bytes, as in the FP32 kernel, and half as
// many memory instructions and address computations are needed per node.
// With one node per thread the 16-bit kernel keeps too few bytes in flight
// to cover HBM latency (its waves spend most cycles waiting on loads while
// the FP32 kernel, issuing the same number of load instructions, moves
// twice the bytes).
//
// Gathers.  Population n of the pair comes from the upstream pair
// (z - ez, z + 1 - ez) of row (x - ex, y - ey): for ez = 0 an aligned 4-byte
// load, for ez = +-1 a 4-byte load at a 2-byte aligned address (the
// hardware and the compiler's HSA target allow unaligned global access).
// The -ez shift goes into the uniform base pointer, so every 32-bit offset
// is a node index in [0, n-2]; an ez = +1 load reads from one element before
// its upstream node (such planes are never the first plane), an ez = -1 load
// up to one element after it (every such plane is followed by another).
//
// Periodic seam.  At z = 0 (ez = +1) and z = nz-2 (ez = -1) the upstream
// pair is split across the periodic seam, (nz-1, 0), and the 4-byte load
// reads one correct half.  The other half comes from a 2-byte patch load
// that goes straight into LDS (global_load_lds_ushort: lane t's value lands
// zero-extended in dword t of a wavefront-uniform LDS address, measured on
// MI355X), so the patches hold no registers while the collision runs; they
// are merged into the pair values after the wait for the plane's loads.
// Held in registers the patches cost some 26 VGPRs per thread.
//
// Alignment.  Tiles start at multiples of 128 nodes, so that a wavefront's
// 256-byte row segment is 256-byte aligned: with the segments shifted by 128
// bytes the same kernel runs 15% slower (measured on MI355X).  This is also
// why the one-node 16-bit kernel, whose wavefronts move 128 bytes per
// instruction, stays further below the copy rate than the FP32 kernel.
//
// The ring pulls are issued after the collision of the plane, not before:
// a ring pull in flight during the collision costs 20+ registers more.
// Ring pairs mostly hit L2 (the neighbour tiles run on the same XCD), and
// only the first wavefronts of a workgroup carry ring pairs.
//
// The per-node arithmetic is lp_prep / lp_out, shared with k_fused_lp.
// This kernel serves tau = 1 (a waiting plane keeps its two momenta only);
// the tau != 1 forms keep 19 values per waiting plane and run the one-node
// kernel, as do walls.
#pragma once

#include "lbm_fused_lp.h"

namespace lbm {
namespace fused {

// Tile of TY rows by TZ node columns; thread t owns pair t % PZ of row t / PZ.
// LDS planes span columns z0-2 .. z0+TZ+1 (pairs on both sides) and rows
// y0-1 .. y0+TY.  Ring pairs: two full rows of TZ/2 + 2 pairs, and the
// pairs (z0-2, z0-1) and (z0+TZ, z0+TZ+1) of each tile row.
template <int TY_, int TZ_> struct PairTile {
    static constexpr int TY = TY_;
    static constexpr int TZ = TZ_;
    static constexpr int PZ = TZ / 2;
    static constexpr int NT = TY * PZ;
    static constexpr int LY = TY + 2;
    static constexpr int LZ = TZ + 4;
    static constexpr int NRING = 2 * (PZ + 2) + 2 * TY;
    static_assert(TZ % 64 == 0, "whole pairs, a wavefront spans whole rows");
    static_assert(NT % 64 == 0 && NT <= 1024, "whole wavefronts");
    static_assert(NRING <= NT, "one ring pair per thread at most");
};

// The 19 stored values of a node pair: v[n] holds node z in its low and
// node z+1 in its high 16 bits.
struct PairVals {
    uint32_t v[19];
};

// Accessor of one node (K = 0: z, 1: z+1) of a pair, in the form
// lp_prep / lp_out / sum19 / HeldJ expect.
template <typename SP, int K> struct PairNode {
    const uint32_t (&v)[19];
    template <int n> __device__ __forceinline__ float get() const
    {
        const unsigned short h = static_cast<unsigned short>(K ? (v[n] >> 16) : v[n]);
        return SP::d(__builtin_bit_cast(typename SP::T, h));
    }
};

// A 4-byte load at a 2-byte aligned address.
typedef uint32_t u32_a2 __attribute__((aligned(2)));

// Base of the loads of population n: its plane, one element earlier for ez
// = +1 when pulling (zsh = 1).
template <typename SP, int n>
__device__ __forceinline__ const typename SP::T* pair_base(const typename SP::T* P, const Grid& g,
                                                           uint32_t zsh)
{
    const typename SP::T* b = P + n * g.ps;
    if constexpr (St<19>::ez[n] == 1) b -= zsh;
    return b;
}

// Pulls the stored values of pair (x, y, z) (wrapped coordinates, z even)
// and its two reference densities R (rin with the node shift, rs
// otherwise).  kind == kHoldsF reads the pair in place.  At the periodic
// seam one half of an ez != 0 pair is wrong until merge_patch().
template <typename SP, bool NS>
__device__ __forceinline__ void pull_pair(const typename SP::T* __restrict__ P,
                                          const float* __restrict__ rin, float rs, const Grid& g,
                                          int kind, int x, int y, int z, PairVals& d,
                                          float (&R)[2])
{
    using S = St<19>;
    const uint32_t nz = static_cast<uint32_t>(g.nz);
    const uint32_t c = (static_cast<uint32_t>(x) * static_cast<uint32_t>(g.ny) +
                        static_cast<uint32_t>(y)) * nz + static_cast<uint32_t>(z);
    if constexpr (NS) {
        const float2 r = *reinterpret_cast<const float2*>(elem(rin, c));
        R[0] = r.x;
        R[1] = r.y;
    } else {
        R[0] = rs;
        R[1] = rs;
    }
    const uint32_t keep = kind == kHoldsF ? 0u : ~0u;
    const NbrOff o(g, x, y, z, keep);
    // The z shift: -1 element (ez = +1) goes into the base, +1 (ez = -1)
    // into the 32-bit offset, which is a node index up <= n-2 before it.  An
    // offset c - 1 would wrap to 2^32 - 1 at node 0 and address 4 GiB past
    // the plane.  pair_base() is shared with the seam patches, so both use
    // one base per population (SGPRs are scarce here).
    const uint32_t zsh = kind == kHoldsF ? 0u : 1u;
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        uint32_t up = c + o.template to<-S::ex[n], -S::ey[n], 0>();
        if constexpr (S::ez[n] == -1) up += zsh;
        d.v[n] = *reinterpret_cast<const u32_a2*>(elem(pair_base<SP, n>(P, g, zsh), up));
    });
}

// D3Q19 populations with ez = +1 and with ez = -1 (St<19>::ez): patch slot
// k holds population kZp[k] of a z = 0 pair or kZm[k] of a z = nz-2 pair.
constexpr int kZp[5] = {5, 9, 11, 16, 18};
constexpr int kZm[5] = {6, 10, 12, 15, 17};

typedef uint32_t __attribute__((address_space(3))) lds_u32;

// Seam patches of pair (x, y, z) into LDS: slot k of this lane is
// slots[k * stride + lane] (slots + k * stride is uniform across the
// wavefront; the lane offset is the hardware's).  seam: 1 = z == 0,
// 2 = z == nz-2, 0 = none (no load).
template <typename SP>
__device__ __forceinline__ void pull_patch(const typename SP::T* __restrict__ P, const Grid& g,
                                           int x, int y, int z, int seam, lds_u32* slots,
                                           int stride)
{
    using S = St<19>;
    if (seam == 0) return;
    const uint32_t nz = static_cast<uint32_t>(g.nz);
    const uint32_t c = (static_cast<uint32_t>(x) * static_cast<uint32_t>(g.ny) +
                        static_cast<uint32_t>(y)) * nz + static_cast<uint32_t>(z);
    const NbrOff o(g, x, y, z);
    // (patches are only loaded when pulling: pair_base() with zsh = 1)
    if (seam == 1) {
        // node z = 0 reads z = nz-1 of its upstream row; the base is one
        // element early
        static_for<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            constexpr int n = kZp[k];
            const uint32_t up = c + o.template to<-S::ex[n], -S::ey[n], 0>() + nz;
            __builtin_amdgcn_global_load_lds(elem(pair_base<SP, n>(P, g, 1u), up), slots + k * stride, 2, 0, 0);
        });
    } else {
        // node z+1 = nz-1 reads z = 0 of its upstream row
        static_for<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            constexpr int n = kZm[k];
            const uint32_t up = c + o.template to<-S::ex[n], -S::ey[n], 0>() - (nz - 2u);
            __builtin_amdgcn_global_load_lds(elem(pair_base<SP, n>(P, g, 1u), up), slots + k * stride, 2, 0, 0);
        });
    }
}

// Folds the patches of this lane (slot k at slots[k * stride + lane]) into
// the pair values; call after the wait for the patch loads.
__device__ __forceinline__ void merge_patch(PairVals& d, int seam, const uint32_t* slots, int stride,
                                            int lane)
{
    if (seam == 1) {
        static_for<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            d.v[kZp[k]] = (d.v[kZp[k]] & 0xffff0000u) | slots[k * stride + lane];
        });
    } else if (seam == 2) {
        static_for<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            d.v[kZm[k]] = (d.v[kZm[k]] & 0xffffu) | (slots[k * stride + lane] << 16);
        });
    }
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left free): the patch loads write
// LDS, not registers, so no register dependence makes the compiler wait.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// A plane waiting one march step (tau = 1): the two momenta.
struct PairHeld {
    HeldJ j[2];
    template <typename SP> __device__ __forceinline__ void pack(const PairVals& l)
    {
        j[0].pack(PairNode<SP, 0>{l.v});
        j[1].pack(PairNode<SP, 1>{l.v});
    }
};

// Collision of a pair from the window (win[i] = {psi, rho} of window node
// i): the arithmetic of lp_prep / lp_out (tau = 1) in the same order, so the
// results are the same bits.  The pair's neighbours e_n are adjacent window
// entries, read in one LDS instruction for psi and rho of both nodes; the
// neighbour densities (the node shift's reference densities) stay in
// registers from here to pair_out.
template <typename SP, bool NS, int LZ>
__device__ __forceinline__ void pair_prep(const PairHeld& d, const int (&sl)[3], int li,
                                          const float2* win, float tau, float G, LpColl (&k)[2],
                                          float (&rn)[2][19])
{
    using S = St<19>;
    const float4 own = *reinterpret_cast<const float4*>(win + sl[1] + li);
    float psi[2][19];
    static_for<19>([&](auto I) {
        constexpr int n = decltype(I)::value;
        if constexpr (n > 0) {
            const int b = sl[1 + S::ex[n]] + li + S::ey[n] * LZ;
            if constexpr (S::ez[n] == 0) {
                const float4 q = *reinterpret_cast<const float4*>(win + b);
                psi[0][n] = q.x; rn[0][n] = q.y; psi[1][n] = q.z; rn[1][n] = q.w;
            } else {
                const float2 qa = win[b + S::ez[n]], qb = win[b + S::ez[n] + 1];
                psi[0][n] = qa.x; rn[0][n] = qa.y; psi[1][n] = qb.x; rn[1][n] = qb.y;
            }
        }
    });
    const float orho[2] = {own.y, own.w}, opsi[2] = {own.x, own.z};
    static_for<2>([&](auto K) {
        constexpr int i = decltype(K)::value;
        LpColl& c = k[i];
        c.rho = orho[i];
        rn[i][0] = orho[i];
        const float jx = d.j[i].jx, jy = d.j[i].jy, jz = d.j[i].jz;
        float gx = 0.f, gy = 0.f, gz = 0.f;
        static_for<19>([&](auto I) {
            constexpr int n = decltype(I)::value;
            if constexpr (n > 0) {
                const float p = static_cast<float>(S::w[n]) * psi[i][n];
                acc_e<S::ex[n]>(gx, p);
                acc_e<S::ey[n]>(gy, p);
                acc_e<S::ez[n]>(gz, p);
            }
        });
        const float ir = __builtin_amdgcn_rcpf(c.rho);
        const float jr = ir * (1.0f / kStoreScale<SP, NS>);
        const float kf = (tau * ir) * ((-G) * opsi[i]);
        c.ue = fmaf(kf, gx, jx * jr);
        c.ve = fmaf(kf, gy, jy * jr);
        c.we = fmaf(kf, gz, jz * jr);
        c.usq15 = 1.5f * fmaf(c.ue, c.ue, fmaf(c.ve, c.ve, c.we * c.we));
        c.t = 0.f;
    });
}

template <typename SP, bool NS, int n>
__device__ __forceinline__ typename SP::T pair_out(const LpColl& k, const float (&rn)[19], float rs)
{
    using S = St<19>;
    constexpr float wn = static_cast<float>(S::w[n]);
    constexpr float sw = kStoreScale<SP, NS> * wn;
    const float ev = edot<S::ex[n], S::ey[n], S::ez[n]>(k.ue, k.ve, k.we);
    const float pm1 = fmaf(fmaf(4.5f, ev, 3.0f), ev, -k.usq15);
    const float rnb = NS ? rn[n] : rs;
    return SP::of_d(fmaf(sw * k.rho, pm1, fmaf(-sw, rnb, sw * k.rho)));
}

// The pair step kernel: one TY x TZ tile per workgroup (z0 a multiple of
// TZ = 128 nodes).
// Occupancy target 2 wavefronts per SIMD: the kernel then takes 130 VGPRs
// (171 with the node shift, whose neighbour densities stay in registers
// through the collision) with no spills.  Held to 128 VGPRs (4 per
// SIMD) the compiler spills 80-100 bytes per thread, and the kernel runs at
// 35 GLUPS instead of 57 (D3Q19 512^3 FP16, MI355X).
template <typename SP, bool NS, typename TL, bool NTS>
__global__ void __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(2)))
    k_fused_pair(const typename SP::T* __restrict__ Pin, int kind, typename SP::T* __restrict__ Pout,
                 Grid g, float tau, float G, float rs, const float* __restrict__ rin,
                 float* __restrict__ rout, int chunk, int order)
{
    constexpr int TY = TL::TY, TZ = TL::TZ, PZ = TL::PZ, LZ = TL::LZ, PL = TL::LY * TL::LZ;
    constexpr int NT = TL::NT;
    constexpr int RW = (TL::NRING + 63) / 64 * 64;   // threads of the ring wavefronts
    // window: {psi, rho} of each node (16-byte aligned: pairs start at even entries)
    __shared__ float4 win4[2 * PL];
    float2* const win = reinterpret_cast<float2*>(win4);
    // seam patch slots: own pair (5 x NT) and ring pair (5 x RW)
    __shared__ uint32_t pown[5 * NT], pring[5 * RW];

    const int nty = g.ny / TY, ntz = g.nz / TZ;
    const int ncx = (g.nx + chunk - 1) / chunk;
    int ty, tz, cx;
    tile_of_block(blockIdx.x, nty, ntz, ncx, order, ty, tz, cx);
    const int y0 = ty * TY, z0 = tz * TZ;
    const int x0 = cx * chunk;
    const int x1 = min(g.nx, x0 + chunk);

    const int tid = threadIdx.x;
    // first thread of this wavefront (uniform, so it lives in an SGPR)
    const int wbase = __builtin_amdgcn_readfirstlane(tid) & ~63;
    const int ly = tid / PZ, lz = 2 * (tid - (tid / PZ) * PZ);
    const int yi = y0 + ly, zi = z0 + lz;
    const int li = (ly + 1) * LZ + (lz + 2);   // window index of node z (z+1 is li+1)
    auto seam_of = [&](int z) { return z == 0 ? 1 : (z == g.nz - 2 ? 2 : 0); };
    // in place (kHoldsF) pairs are read whole: no patches
    const bool pulled = kind != kHoldsF;
    const int seam = pulled ? seam_of(zi) : 0;
    // wavefront-uniform LDS bases of this thread's patch slots
    lds_u32* const sown = (lds_u32*)(pown + wbase);
    lds_u32* const sring = (lds_u32*)(pring + wbase);

    // this thread's ring pair, if any
    const bool has_ring = tid < TL::NRING;
    int ry = 0, rz = 0, lr = 0;
    if (has_ring) {
        int row, col;
        if (tid < PZ + 2) { row = 0; col = 2 * tid; }
        else if (tid < 2 * (PZ + 2)) { row = TL::LY - 1; col = 2 * (tid - (PZ + 2)); }
        else if (tid < 2 * (PZ + 2) + TY) { row = 1 + (tid - 2 * (PZ + 2)); col = 0; }
        else { row = 1 + (tid - 2 * (PZ + 2) - TY); col = LZ - 2; }
        ry = wrap(y0 + row - 1, g.ny);
        rz = wrap(z0 + col - 2, g.nz);
        lr = row * LZ + col;
    }
    const int rseam = (pulled && has_ring) ? seam_of(rz) : 0;

    auto slot = [&](int x) { return ((x - x0 + 1) & 3) * PL; };
    auto xw = [&](int x) { return x >= g.nx ? x - g.nx : (x < 0 ? x + g.nx : x); };
    const uint32_t plane = static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(g.nz);
    const uint32_t cyz = static_cast<uint32_t>(yi) * static_cast<uint32_t>(g.nz) +
                         static_cast<uint32_t>(zi);
    constexpr float isc = 1.0f / kStoreScale<SP, NS>;

    // pulls (values in registers, seam patches into LDS)
    auto pull_own = [&](int xp, PairVals& d, float (&R)[2]) {
        pull_pair<SP, NS>(Pin, rin, rs, g, kind, xp, yi, zi, d, R);
        pull_patch<SP>(Pin, g, xp, yi, zi, seam, sown, NT);
    };
    auto pull_ring = [&](int xp, PairVals& d, float (&R)[2]) {
        if (has_ring) {
            pull_pair<SP, NS>(Pin, rin, rs, g, kind, xp, ry, rz, d, R);
            pull_patch<SP>(Pin, g, xp, ry, rz, rseam, sring, RW);
        }
    };
    // window entries of a pair at window indices at, at+1; returns the densities
    auto put_win = [&](int at, const PairVals& d, const float (&R)[2], float (&rho)[2]) {
        rho[0] = fmaf(sum19(PairNode<SP, 0>{d.v}), isc, R[0]);
        rho[1] = fmaf(sum19(PairNode<SP, 1>{d.v}), isc, R[1]);
        *reinterpret_cast<float4*>(win + at) = make_float4(psi_lp(rho[0]), rho[0], psi_lp(rho[1]), rho[1]);
    };
    // after wait_vmem(): merge the patches, fill the window; the density
    // field is written for this thread's pairs of the planes of this chunk
    auto ring_win = [&](int s, PairVals& d, const float (&R)[2]) {
        if (has_ring) {
            float rr[2];
            merge_patch(d, rseam, pring, RW, tid);
            put_win(s + lr, d, R, rr);
        }
    };
    auto own_win = [&](int xu, PairVals& d, const float (&R)[2]) {
        float rho[2];
        merge_patch(d, seam, pown, NT, tid);
        put_win(slot(xu) + li, d, R, rho);
        if constexpr (NS) {
            if (xu >= x0 && xu < x1)
                *reinterpret_cast<float2*>(elem(rout, static_cast<uint32_t>(xu) * plane + cyz)) =
                    make_float2(rho[0], rho[1]);
        }
    };

    PairHeld fa, fb, fc;
    float Ra[2], Rb[2], Rc[2];
    // prologue: window planes x0-1 .. x0+1, values of x0 (fa) and x0+1 (fb);
    // each pull completes before the next one starts
    auto prologue = [&](int xu, float (&R)[2], PairHeld* held) {
        {
            PairVals r;
            float Rr[2];
            pull_ring(xw(xu), r, Rr);
            wait_vmem();
            ring_win(slot(xu), r, Rr);
        }
        PairVals l;
        pull_own(xw(xu), l, R);
        wait_vmem();
        own_win(xu, l, R);
        if (held) held->template pack<SP>(l);
        // the patch slots are read before the next pull refills them
        __syncthreads();
    };
    prologue(x0 - 1, Rc, nullptr);
    prologue(x0, Ra, &fa);
    prologue(x0 + 1, Rb, &fb);

    auto collide = [&](int x, const PairHeld& d, const float (&R)[2]) {
        const int sl[3] = {slot(x - 1), slot(x), slot(x + 1)};
        LpColl k[2];
        float rn[2][19];
        pair_prep<SP, NS, LZ>(d, sl, li, win, tau, G, k, rn);
        const uint32_t c = static_cast<uint32_t>(x) * plane + cyz;
        static_for<19>([&](auto I) {
            constexpr int n = decltype(I)::value;
            const typename SP::T v0 = pair_out<SP, NS, n>(k[0], rn[0], rs);
            const typename SP::T v1 = pair_out<SP, NS, n>(k[1], rn[1], rs);
            const uint32_t pv = static_cast<uint32_t>(__builtin_bit_cast(unsigned short, v0)) |
                                (static_cast<uint32_t>(__builtin_bit_cast(unsigned short, v1)) << 16);
            uint32_t* base = reinterpret_cast<uint32_t*>(elem(Pout + n * g.ps, c));
            if constexpr (NTS) __builtin_nontemporal_store(pv, base);
            else *base = pv;
        });
    };

    auto step = [&](int x, const PairHeld& fcur, const float (&Rcur)[2], PairHeld& fnn,
                    float (&Rnn)[2]) {
        const int x2 = xw(x + 2);
        PairVals ln, lr2;
        float Rr[2];
        // own pulls of plane x+2 in flight during the collision of plane x;
        // the ring pulls follow the collision (fewer registers held across it)
        pull_own(x2, ln, Rnn);
        collide(x, fcur, Rcur);
        pull_ring(x2, lr2, Rr);
        wait_vmem();
        ring_win(slot(x + 2), lr2, Rr);
        own_win(x + 2, ln, Rnn);
        fnn.template pack<SP>(ln);
        __syncthreads();
    };
    int x = x0;
    for (; x + 2 < x1; x += 3) {
        step(x, fa, Ra, fc, Rc);
        step(x + 1, fb, Rb, fa, Ra);
        step(x + 2, fc, Rc, fb, Rb);
    }
    if (x < x1) step(x, fa, Ra, fc, Rc);
    if (x + 1 < x1) step(x + 1, fb, Rb, fa, Ra);
}

}  // namespace fused
}  // namespace lbm
