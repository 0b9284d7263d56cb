// tile_map.hpp — how the fused step's tiles map coordinates outside the
// domain back into it (host and device, so the host can check the mapping).
#pragma once

namespace lbm {

// Map a coordinate of the extended tile into the domain: periodic side wraps,
// any other side clamps (the rule of nb_index for arbitrary distance).
// map_coord serves z planes: the kernels request planes z0 - 2 .. z1 + 1 of a
// chunk [z0, z1), at most two cells below a face and one above it (one
// addition or subtraction of n, or two for n = 1). map_col serves the x and y
// columns, fixed per thread: the inactive columns of a tile wider than the
// domain can lie further out and wrap by a remainder, so that every load
// stays inside the domain.
__host__ __device__ __forceinline__ int map_coord(int t, int n, bool lo_per, bool hi_per) {
    if (t < 0) {
        if (!lo_per) return 0;
        t += n;
        return t < 0 ? t + n : t;
    }
    if (t >= n) {
        if (!hi_per) return n - 1;
        t -= n;
        return t >= n ? t - n : t;
    }
    return t;
}
__host__ __device__ __forceinline__ int map_col(int t, int n, bool lo_per, bool hi_per) {
    if (t < 0) {
        if (!lo_per) return 0;
        t += n;
        return t >= 0 ? t : (n - 1) - ((-t - 1) % n);
    }
    if (t >= n) {
        if (!hi_per) return n - 1;
        t -= n;
        return t < n ? t : t % n;
    }
    return t;
}

}  // namespace lbm
