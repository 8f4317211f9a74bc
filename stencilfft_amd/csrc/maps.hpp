// maps.hpp -- the reference's layout algebra (zmodel/utils.h:99-320) as
// __host__ __device__ functions of a runtime shape.  The C ABI's host index
// maps and every kernel that needs a global coordinate use these.
#pragma once

#include <hip/hip_runtime.h>

namespace zm {

struct MapShape {
    int n, p, b;
    int l, vec, vp, m;
};

__host__ __device__ inline MapShape make_map_shape(int n, int p, int b) {
    MapShape s;
    s.n = n;
    s.p = p;
    s.b = b;
    s.l = n / p;
    s.vec = s.l / b;
    s.vp = s.vec / p;
    s.m = b * p;
    return s;
}

// zmodel/utils.h:122-124
__host__ __device__ inline int packed_to_block_cyclic(const MapShape& s, int packed_col) {
    return (packed_col / s.vec) + (packed_col % s.vec) * s.b;
}
// zmodel/utils.h:126-128
__host__ __device__ inline int block_cyclic_to_packed(const MapShape& s, int block_cyclic_col) {
    return (block_cyclic_col % s.b) * s.vec + (block_cyclic_col / s.b);
}

// zmodel/utils.h:166-170: a local row / packed column of rank (rid, cid) to
// the global grid
__host__ __device__ inline int local_row_to_global(const MapShape& s, int rank_coord, int lr) {
    return (lr / s.b) * s.b * s.p + rank_coord * s.b + (lr % s.b);
}
__host__ __device__ inline int packed_col_to_global(const MapShape& s, int rank_coord, int lc) {
    const int bcc = packed_to_block_cyclic(s, lc);
    return (bcc / s.b) * s.b * s.p + rank_coord * s.b + (bcc % s.b);
}
// zmodel/utils.h:176-185
__host__ __device__ inline void global_row_to_local(const MapShape& s, int gr, int* rank_coord, int* lr) {
    *rank_coord = (gr / s.b) % s.p;
    *lr = (gr / (s.b * s.p)) * s.b + (gr % s.b);
}
__host__ __device__ inline void global_col_to_local(const MapShape& s, int gc, int* rank_coord, int* lc) {
    *rank_coord = (gc / s.b) % s.p;
    const int bcc = (gc / (s.b * s.p)) * s.b + (gc % s.b);
    *lc = block_cyclic_to_packed(s, bcc);
}

// zmodel/utils.h:232-252
__host__ __device__ inline int lp(const MapShape& s, int i) { return (i / s.vp) + (i % s.vp) * s.m; }
__host__ __device__ inline int ct(const MapShape& s, int p) { return (p % s.m) * s.vec + (p / s.m); }
__host__ __device__ inline int lp_inv(const MapShape& s, int j) { return (j % s.m) * s.vp + (j / s.m); }
__host__ __device__ inline int ct_inv(const MapShape& s, int k) { return (k % s.vec) * s.m + (k / s.vec); }

// One reciprocal-space axis: the frequency held by local index `local` of a
// rank at coordinate `rank_coord` along the axis that index went through
// (zmodel/utils.h:265-268, 298-302), and back (utils.h:274-282, 308-320).
__host__ __device__ inline int recip_to_freq(const MapShape& s, int rank_coord, int local) {
    return ct(s, rank_coord * s.l + lp(s, local));
}
__host__ __device__ inline void freq_to_recip(const MapShape& s, int k, int* rank_coord, int* local) {
    const int pk = ct_inv(s, k);
    *rank_coord = pk / s.l;
    *local = lp_inv(s, pk % s.l);
}

}  // namespace zm
