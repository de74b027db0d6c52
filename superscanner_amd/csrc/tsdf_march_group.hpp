// tsdf_march_group.hpp -- the raycast march of one ray carried by a group of
// G lanes of a wavefront (G = 16: four rays per wavefront), for the rays
// that the per-lane march (march, tsdf_march.hpp) deferred after its round
// budget: the long rays that graze surfaces through many bricks.
//
// A round of the group evaluates G consecutive positions of its ray at once,
// lane t taking the position t steps ahead (reached by the same sequence of
// float additions the reference makes), and then applies the reference's
// decisions to them in order: the first position where the reference would
// stop decides, everything before it is taken as the reference takes it.
// The walk continues across brick boundaries within a round: a position in
// another block looks that block up (the lookup the reference's outer loop
// makes there) and, when it hits, samples the new brick; a miss or an
// unfinished probe hands the position back to the outer loop.  Every
// lookup, sample and floating-point operation, and every count, is the
// reference march's (GPUVoxelGrid.cpp:589-707), in its order.
//
// Every branch below depends only on values the whole group shares, so the
// group's lanes stay together; cross-lane operations read lanes of the
// group only.
#pragma once

#include "tsdf_march.hpp"

#pragma clang fp contract(off)

namespace tsdf {

template <int G>
struct Group {
    static_assert(G == 8 || G == 16 || G == 32, "a group is 8, 16 or 32 consecutive lanes");
    int t;     // lane within the group
    int base;  // the group's first lane in the wavefront
    __device__ Group()
    {
        const int lane = __lane_id();
        t = lane & (G - 1);
        base = lane & ~(G - 1);
    }
    // bit k: predicate of lane k of the group
    __device__ unsigned int ballot(bool p) const
    {
        return (unsigned int)((__ballot(p) >> base) & ((1ull << G) - 1ull));
    }
    __device__ int from(int v, int k) const { return __shfl(v, base + k); }
    __device__ unsigned int from(unsigned int v, int k) const { return (unsigned int)__shfl((int)v, base + k); }
    __device__ float from(float v, int k) const { return __shfl(v, base + k); }
    __device__ unsigned long long from(unsigned long long v, int k) const { return __shfl(v, base + k); }
    __device__ f3 from(f3 v, int k) const { return {from(v.x, k), from(v.y, k), from(v.z, k)}; }
};

// probe_run (tsdf_march.hpp) with the group: each round looks up G
// positions at once -- lane t the one after t (first round) or t + 1 steps.
template <int G>
__device__ int gprobe_run(const GridView& g, const RayParams& rp, const Group<G>& gr, f3& rv, float& l, const f3 delta,
                          float dl, int& steps, unsigned int& probes, float l_safe, bool pre_valid, uint32_t pre_key,
                          uint32_t pre_bk, unsigned long long pre_slot)
{
    const float dmax = rp.dmax;
    int id = -1;
    bool first = true;
    if (pre_valid) {
        const int r0 = first_slot(g, pre_key, pre_slot);
        ++probes;
        id = r0 == kProbeMore ? probe_from(g, pre_key, (pre_bk + 1) & g.mask) : r0;
        first = false;
        if (id >= 0)
            return id;
    }
    while (first || (id < 0 && l < dmax)) {
        // tile depth bounds, as in probe_run
        if (first && l < l_safe) {
            ++probes;
            first = false;
        }
        if (!first) {
            for (;;) {
                const float ln = l + dl;
                if (!(ln < l_safe))
                    break;
                rv = add3(rv, delta);
                l = ln;
                ++steps;
                ++probes;
            }
        }
        const int nst = first ? gr.t : gr.t + 1;
        f3 r = rv;
        float lr = l;
        for (int k = 1; k <= G; ++k)
            if (k <= nst) {
                r = add3(r, delta);
                lr = lr + dl;
            }
        int x, y, z;
        vb3(r, rp, x, y, z);
        const uint32_t key = prime_hash(x, y, z), bk = mix32(key) & g.mask;
        const int r0 = first_slot(g, key, g.slots[bk]);
        // lookup k is made when every earlier one missed with l < d_max after it
        const unsigned int m = gr.ballot(!(r0 == -1 && lr < dmax));
        const int s = m ? __ffs(m) - 1 : G - 1;
        probes += m ? (unsigned int)(s + 1) : (unsigned int)G;
        steps += first ? s : s + 1;
        rv = gr.from(r, s);
        l = gr.from(lr, s);
        first = false;
        id = -1;
        if (m) {
            id = gr.from(r0, s);
            if (id == kProbeMore)
                id = probe_from(g, gr.from(key, s), (gr.from(bk, s) + 1u) & g.mask);
        }
    }
    return id;
}

// The voxel walk from rv (position 0, in block (bx, by, bz) = brick cur),
// G positions a round, across bricks.  Returns true when the ray is done
// (out set), false when the outer loop continues at rv with the lookup
// there loaded ahead (pre_*).
template <typename T, bool kColor, int G>
__device__ bool gwalk(const GridView& g, const RayParams& rp, const Group<G>& gr, f3& rv, float& l, const f3 dir,
                      const f3 step, int bx, int by, int bz, int cur, bool& Lp, bool& pre_valid, uint32_t& pre_key,
                      uint32_t& pre_bk, unsigned long long& pre_slot, RayResult& out)
{
    const float dmax = rp.dmax, vl = rp.vl;
    for (;;) {
        f3 r = rv;
        float lr = l;
        for (int k = 1; k < G; ++k)
            if (k <= gr.t) {
                r = add3(r, step);
                lr = lr + vl;
            }
        int qx, qy, qz, wx, wy, wz;
        vb3(r, rp, qx, qy, qz);
        vin3(r, qx, qy, qz, rp, wx, wy, wz);
        const int vk = vin_ok(wx, wy, wz) ? wx + kVPE * wy + kVPE * kVPE * wz : -1;
        // the block of the position before (the walk's block for position 0)
        const int pl = gr.t > 0 ? gr.t - 1 : 0;
        int px = gr.from(qx, pl), py = gr.from(qy, pl), pz = gr.from(qz, pl);
        if (gr.t == 0) {
            px = bx;
            py = by;
            pz = bz;
        }
        const bool trans = qx != px || qy != py || qz != pz;
        const bool lok = lr < dmax;
        const bool inwalk = qx == bx && qy == by && qz == bz;
        // samples of the walk's brick and the lookups of the other blocks ...
        uint32_t key = 0, bk = 0;
        unsigned long long sl = 0;
        float s = 0.f, w = 0.f;
        if (!inwalk) {
            key = prime_hash(qx, qy, qz);
            bk = mix32(key) & g.mask;
            sl = g.slots[bk];
        } else if (lok && vk >= 0) {
            ld_pair((const T*)g.payload + (size_t)cur * (2 * kVPB), vk, s, w);
        }
        const int id = inwalk ? cur : first_slot(g, key, sl);
        // ... then the samples of the blocks whose lookup hit
        if (!inwalk && id >= 0 && lok && vk >= 0)
            ld_pair((const T*)g.payload + (size_t)id * (2 * kVPB), vk, s, w);
        const bool wpos = w > 0, spos = s > 0;
        // the reference's decision at this position, if it gets here:
        // 1 the outer loop ends (l >= d_max), 2 the outer loop takes over
        // (the lookup of a new block missed or needs more probing), 3 voxel
        // index out of the brick, 4 first sample at or behind the surface
        const int evt = !lok ? 1 : (trans && id < 0) ? 2 : vk < 0 ? 3 : (wpos && !spos) ? 4 : 0;
        const unsigned int m = gr.ballot(evt != 0);
        const int e = m ? __ffs(m) - 1 : G;
        // the lookups of the blocks entered up to e (the one at e when it hit)
        const unsigned int hits = gr.ballot(trans && lok && id >= 0);
        const unsigned int upto = e >= G ? (unsigned int)((1ull << G) - 1ull) : ((2u << e) - 1u);
        out.probes += (unsigned int)__popc(hits & upto);
        out.samples += (unsigned int)e + ((e < G && gr.from(evt, e) == 4) ? 1u : 0u);
        if (e > 0)
            Lp = gr.from((int)wpos, e - 1) != 0;
        if (e == G) {
            // every position taken: continue from the one after the last
            rv = add3(gr.from(r, G - 1), step);
            l = gr.from(lr, G - 1) + vl;
            bx = gr.from(qx, G - 1);
            by = gr.from(qy, G - 1);
            bz = gr.from(qz, G - 1);
            cur = gr.from(id, G - 1);
            continue;
        }
        const int ev = gr.from(evt, e);
        if (ev == 1)
            return true;  // miss
        rv = gr.from(r, e);
        l = gr.from(lr, e);
        if (ev == 2) {
            pre_valid = true;
            pre_key = gr.from(key, e);
            pre_bk = gr.from(bk, e);
            pre_slot = gr.from(sl, e);
            return false;
        }
        if (ev == 3) {
            out.code = -kRayUbVoxelIndex;
            return true;
        }
        if (!Lp)
            return true;  // Fn: miss
        crossing<T, kColor, false>(g, rp, Footprint{}, rv, l, dir, gr.from(qx, e), gr.from(qy, e), gr.from(qz, e),
                                   gr.from(id, e), gr.from(s, e), gr.from(w, e), gr.from(vk, e), out);
        return true;
    }
}

// The march of a deferred ray (RayState) to its end, by the group.
template <typename T, bool kColor, int G>
__device__ RayResult march_group(const GridView& g, const Pose& pose, const RayParams& rp, const RayState& st,
                                 const Group<G>& gr)
{
    RayResult out;
    out.p = {0.f, 0.f, 0.f};
    out.c = {0.f, 0.f, 0.f};
    out.code = 0;
    out.samples = st.samples;
    out.probes = st.probes;
    const float vl = rp.vl, vbl = rp.vbl, dt = rp.dt, dmax = rp.dmax;
    const int i = (int)(st.pix / (uint32_t)rp.W), j = (int)(st.pix % (uint32_t)rp.W);
    f3 n, rv0, dir;
    ray_setup(g, pose, i, j, n, rv0, dir);
    const float l_safe = (rp.use_tb && st.zt >= 0.0f) ? tb_l_safe(st.zt, n.z, dmax) : -1.0f;
    const f3 skip = {vbl * dir.x, vbl * dir.y, vbl * dir.z};
    const f3 step = {vl * dir.x, vl * dir.y, vl * dir.z};
    f3 rv = {st.rx, st.ry, st.rz};
    float l = st.l;
    bool Lp = (st.flags & 1u) != 0;
    bool pre_valid = (st.flags & 2u) != 0;
    uint32_t pre_key = st.pre_key, pre_bk = st.pre_bk;
    unsigned long long pre_slot = st.pre_slot;
    while (l < dmax) {
        int count = 0;
        int id = gprobe_run<G>(g, rp, gr, rv, l, skip, vbl, count, out.probes, l_safe, pre_valid, pre_key, pre_bk,
                               pre_slot);
        pre_valid = false;
        if (!(l < dmax))
            break;
        if (count > 0) {
            const float back = rp.back;
            rv.x -= back * dir.x;
            rv.y -= back * dir.y;
            rv.z -= back * dir.z;
            l -= back;
            const f3 dstep = {dt * dir.x, dt * dir.y, dt * dir.z};
            int steps = 0;
            id = gprobe_run<G>(g, rp, gr, rv, l, dstep, dt, steps, out.probes, l_safe, false, 0u, 0u, 0ull);
        }
        if (!(l < dmax))
            break;
        int bx, by, bz;
        vb3(rv, rp, bx, by, bz);
        if (gwalk<T, kColor, G>(g, rp, gr, rv, l, dir, step, bx, by, bz, id, Lp, pre_valid, pre_key, pre_bk,
                                pre_slot, out))
            return out;
    }
    return out;
}

}  // namespace tsdf
