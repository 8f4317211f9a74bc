// Host-only check of the C++ facade's HDF5 wrappers (no GPU calls): the
// reference's init_host ramp (utils.cu:9-81), stored and loaded back, and
// read through a second distribution of the same field.
#include <cstdio>
#include <vector>

#include "zmodel.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const char* file = argv[1];
    const zm_shape a{2, 64, 4, 2}, b{2, 64, 2, 8};
    zm_dims da, db;
    if (zm_shape_check(&a, &da) != ZM_OK || zm_shape_check(&b, &db) != ZM_OK) return 3;
    std::vector<zmodel::Complex> g(da.field), back(da.field), other(db.field), want(db.field);
    zmodel::init_host(a, g.data());
    zmodel::store_host_hdf5(a, g.data(), file, "ramp");
    zmodel::init_host_hdf5(a, back.data(), file, "ramp");
    for (size_t i = 0; i < g.size(); i++)
        if (back[i].x != g[i].x || back[i].y != g[i].y) return 4;
    // the ramp holds each point's global index, so any distribution reads it back as its own ramp
    zmodel::init_host_hdf5(b, other.data(), file, "ramp");
    zmodel::init_host(b, want.data());
    for (size_t i = 0; i < want.size(); i++)
        if (other[i].x != want[i].x || other[i].y != want[i].y) return 5;
    try {
        zmodel::init_host_hdf5(a, back.data(), file, "missing");
        return 6;
    } catch (const zmodel::Error& e) {
        if (e.status != ZM_ERR_IO) return 7;
    }
    std::printf("facade hdf5 ok\n");
    return 0;
}
