// hdf5_example.cpp -- a field through the HDF5 file format and the GPU: the
// reference's init_host ramp (zmodel/utils.cu:9-81) is stored with
// store_host_hdf5, read back with init_host_hdf5 (utils.cu:103-281), copied to
// the device, run through Laplacian -> forward FFT -> inverse FFT (1/N^d), and
// the result is stored and read back again, all through the zmodel.hpp facade.
//
// usage: zmodel_hdf5_example FILE [dims N P B]
// Checks that the file round trips are bit-exact and that the GPU result
// matches the ramp's periodic Laplacian (zero except across the wrap-around)
// computed on the host from the global coordinates; exits non-zero otherwise.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "zmodel.hpp"

using namespace zmodel;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s FILE [dims N P B]\n", argv[0]);
        return 2;
    }
    const char* file = argv[1];
    zm_shape shape{2, 256, 4, 8};
    if (argc >= 6) shape = zm_shape{atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5])};
    try {
        FftHostPlans plans;
        init_plans(plans, shape, ZM_FP64);
        FftBuffers buffers;
        init_buffers<true>(buffers, plans);
        const size_t total = (size_t)plans.dims.field;
        const int N = shape.n, d3 = shape.dims == 3;

        // host ramp -> file -> host: the reference's input path
        std::vector<Complex> ramp(total), loaded(total), result(total), reloaded(total);
        init_host(shape, ramp.data());
        store_host_hdf5(shape, ramp.data(), file, "ramp");
        init_host_hdf5(shape, loaded.data(), file, "ramp");
        for (size_t i = 0; i < total; i++)
            if (loaded[i].x != ramp[i].x || loaded[i].y != ramp[i].y) {
                std::fprintf(stderr, "ramp differs after the file round trip at %zu\n", i);
                return 1;
            }

        // device: Laplacian, forward, inverse
        check_hip(hipMemcpy(buffers.device_buf0, loaded.data(), buffers.bytes, hipMemcpyHostToDevice), "H2D");
        laplace(plans, buffers.device_buf0, buffers.device_buf1);
        check_hip(hipMemcpy(buffers.device_buf0, buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToDevice), "D2D");
        const double s = 1.0 / std::pow(double(N), shape.dims);
        if (d3) {
            forward_fft_3d<true>(buffers, plans);
            check_hip(hipMemcpy(buffers.device_buf0, buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToDevice), "D2D");
            inverse_fft_3d<true>(buffers, plans, Complex{s, 0.0});
        } else {
            forward_fft<true>(buffers, plans);
            check_hip(hipMemcpy(buffers.device_buf0, buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToDevice), "D2D");
            inverse_fft<true>(buffers, plans, Complex{s, 0.0});
        }
        check_hip(hipMemcpy(result.data(), buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToHost), "D2H");

        // device result -> file -> host, bit for bit
        store_host_hdf5(shape, result.data(), file, "result");
        init_host_hdf5(shape, reloaded.data(), file, "result");
        for (size_t i = 0; i < total; i++)
            if (reloaded[i].x != result[i].x || reloaded[i].y != result[i].y) {
                std::fprintf(stderr, "result differs after the file round trip at %zu\n", i);
                return 1;
            }

        // against the periodic Laplacian of the ramp x = (gd N + gr) N + gc
        const int P = shape.p, L = plans.dims.l;
        auto v = [&](int z, int r, int c) {
            return (double(((z + N) % N)) * N + double((r + N) % N)) * N + double((c + N) % N);
        };
        double max_err = 0.0, max_ref = 0.0;
        const int nd = d3 ? P : 1, nl = d3 ? L : 1;
        for (int did = 0; did < nd; did++)
            for (int rid = 0; rid < P; rid++)
                for (int cid = 0; cid < P; cid++)
                    for (int ld = 0; ld < nl; ld++)
                        for (int lr = 0; lr < L; lr++)
                            for (int lc = 0; lc < L; lc++) {
                                int gd = 0, gr, gc;
                                if (d3) real_to_global_3d(shape, rid, cid, did, ld, lr, lc, gd, gr, gc);
                                else real_to_global(shape, rid, cid, lr, lc, gr, gc);
                                double ref = v(gd, gr - 1, gc) + v(gd, gr + 1, gc) + v(gd, gr, gc - 1) +
                                             v(gd, gr, gc + 1) - (d3 ? 6 : 4) * v(gd, gr, gc);
                                if (d3) ref += v(gd - 1, gr, gc) + v(gd + 1, gr, gc);
                                const size_t at = rank_offset(plans, rid, cid, did) + ((size_t)ld * L + lr) * L + lc;
                                max_err = std::fmax(max_err, std::hypot(reloaded[at].x - ref, reloaded[at].y));
                                max_ref = std::fmax(max_ref, std::fabs(ref));
                            }
        std::printf("%dD N=%d P=%d B=%d: file round trips bit-exact; Laplacian -> FFT -> IFFT max error %.3e "
                    "(max |laplace| %.3e)\n", shape.dims, N, shape.p, shape.b, max_err, max_ref);
        destroy_buffers<true>(buffers);
        destroy_plans(plans);
        const bool ok = max_err <= 1e-12 * max_ref;
        std::printf("%s\n", ok ? "OK" : "FAILED");
        return ok ? 0 : 1;
    } catch (const Error& e) {
        std::fprintf(stderr, "zmodel error: %s\n", e.what());
        return 1;
    }
}
