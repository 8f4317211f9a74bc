// example.cpp -- the reference's end-to-end example() (zmodel/main.cu:8-102)
// written against the zmodel.hpp facade: ramp field -> 5-point Laplacian ->
// forward FFT -> inverse FFT (1/N^2), all P x P virtual ranks on one GPU.
//
// usage: zmodel_example [N P B] [fp64|fp32]
// Prints the round-trip error against the Laplacian computed on the host from
// the global coordinates (the ramp is linear, so its periodic Laplacian is zero
// except across the wrap-around) and exits non-zero if it is too large.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "zmodel.hpp"

using namespace zmodel;

int main(int argc, char** argv) {
    zm_shape shape{2, 1024, 2, 16};
    if (argc >= 4) shape = zm_shape{2, atoi(argv[1]), atoi(argv[2]), atoi(argv[3])};
    if (argc >= 5 && std::strcmp(argv[4], "fp64") != 0) {
        std::fprintf(stderr, "this example runs the reference's FP64 path only\n");
        return 2;
    }
    try {
        FftHostPlans plans;
        init_plans(plans, shape, ZM_FP64);
        FftBuffers buffers;
        init_buffers<true>(buffers, plans);
        LaplaceBuffers laplace_buffers;
        init_laplace_buffers(laplace_buffers);
        const int N = shape.n;
        const size_t total = (size_t)plans.dims.field;

        // init_host + H2D (main.cu:46-48)
        Complex* host_in = static_cast<Complex*>(buffers.host_buf0);
        init_host(shape, host_in);
        check_hip(hipMemcpy(buffers.device_buf0, host_in, buffers.bytes, hipMemcpyHostToDevice), "H2D");

        // Laplace routine (main.cu:50-58): pack / communicate are direct slab reads here
        pack_laplace(buffers.device_buf0, laplace_buffers);
        communicate_laplace(laplace_buffers);
        laplace(plans, buffers.device_buf0, buffers.device_buf1);
        check_hip(hipMemcpy(buffers.device_buf0, buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToDevice), "D2D");

        // forward, then inverse with 1/N^2 (main.cu:64-80)
        forward_fft<true>(buffers, plans);
        check_hip(hipMemcpy(buffers.device_buf0, buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToDevice), "D2D");
        const double s = 1.0 / (double(N) * double(N));
        inverse_fft<true>(buffers, plans, Complex{s, 0.0});
        Complex* host_out = static_cast<Complex*>(buffers.host_buf1);
        check_hip(hipMemcpy(host_out, buffers.device_buf1, buffers.bytes, hipMemcpyDeviceToHost), "D2H");

        // expected: periodic 5-point Laplacian of x[gr][gc] = gr*N + gc, mapped
        // through real_to_global
        double max_err = 0.0, max_ref = 0.0;
        const int P = shape.p, L = plans.dims.l;
        for (int rid = 0; rid < P; rid++)
            for (int cid = 0; cid < P; cid++)
                for (int lr = 0; lr < L; lr++)
                    for (int lc = 0; lc < L; lc++) {
                        int gr, gc;
                        real_to_global(shape, rid, cid, lr, lc, gr, gc);
                        auto v = [&](int r, int c) { return double(((r + N) % N) * N + (c + N) % N); };
                        const double ref = v(gr - 1, gc) + v(gr + 1, gc) + v(gr, gc - 1) + v(gr, gc + 1) - 4 * v(gr, gc);
                        const Complex got = host_out[rank_offset(plans, rid, cid) + (size_t)lr * L + lc];
                        max_err = std::fmax(max_err, std::hypot(got.x - ref, got.y));
                        max_ref = std::fmax(max_ref, std::fabs(ref));
                    }
        std::printf("N_DIM: %d, B_DIM: %d P_DIM: %d (%zu elements, %d virtual ranks)\n", N, shape.b, shape.p, total,
                    plans.dims.nranks);
        std::printf("round trip max error %.3e (max |laplace| %.3e)\n", max_err, max_ref);
        destroy_laplace_buffers(laplace_buffers);
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
