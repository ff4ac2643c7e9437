// example_gpu.cpp -- the C++ drop-in interface in use, as the reference's
// src/examples/example_gpu.cpp:31-104 uses it, with a synthetic Sandia-style
// record instead of a MAT file:
//   1. sph_sar_data_callback_gpu(...) on raw arrays -> BMP file;
//   2. SAR_Aperture + SAR_ImageFormationParameters with zero padding and an
//      explicit pixel grid -> cuda_focus_SAR_image -> brightest pixel, which
//      must land on the strongest synthetic target;
//   3. optionally (argv[2], argv[3]) that image as raw complex64 (x-major)
//      and as a BMP from writeBMPFile.
// Prints one line "peak x=<m> y=<m> |I|=<value>" and exits 0 on success.
#include <charlotte_sar_api.hpp>

#include <cmath>
#include <cstdio>
#include <vector>

int main(int argc, char **argv) {
    outputfile = argc > 1 ? argv[1] : "example_gpu.bmp";
    const int K = 256, Np = 128;
    const double f0 = 9.6e9, bw = 640e6, adf = 1.0e6;  // A/D rate (arbitrary units)
    const double df = bw / (K - 1), chirp = df * adf;   // chirpRate / ADF = deltaF
    const double tx[3] = {0.0, 1.5, -2.0}, ty[3] = {0.0, -1.0, 2.5}, amp[3] = {1.0, 0.6, 0.4};

    std::vector<float> sample(2 * K * Np), startF(Np), chirpRate(Np), x(Np), y(Np), z(Np);
    for (int n = 0; n < Np; n++) {
        const double th = (90.0 - 2.0 + 4.0 * n / (Np - 1)) * PI / 180.0;
        const double ax = 10000.0 * std::cos(th), ay = 10000.0 * std::sin(th), az = 7500.0;
        x[n] = (float)ax;
        y[n] = (float)ay;
        z[n] = (float)az;
        startF[n] = (float)f0;
        chirpRate[n] = (float)chirp;
        // R0 = |a_n| of the float positions, as initialize_SAR_Aperture_Data derives it
        const double r0 = std::sqrt((double)x[n] * x[n] + (double)y[n] * y[n] + (double)z[n] * z[n]);
        for (int k = 0; k < K; k++) {
            double re = 0, im = 0;
            for (int t = 0; t < 3; t++) {
                const double dx = x[n] - tx[t], dy = y[n] - ty[t], dz = z[n];
                const double dR = std::sqrt(dx * dx + dy * dy + dz * dz) - r0;
                const double ph = -4.0 * PI * (f0 + k * df) * dR / CLIGHT;
                re += amp[t] * std::cos(ph);
                im += amp[t] * std::sin(ph);
            }
            sample[2 * ((size_t)n * K + k)] = (float)re;
            sample[2 * ((size_t)n * K + k) + 1] = (float)im;
        }
    }

    // 1. the callback API (default image parameters, BMP output)
    if (sph_sar_data_callback_gpu<float>(0, (float)adf, sample.data(), K, startF.data(),
                                         chirpRate.data(), x.data(), y.data(), z.data(),
                                         Np) != EXIT_SUCCESS) {
        std::fprintf(stderr, "callback failed\n");
        return 1;
    }

    // 2. the focusing call with an explicit, zero-padded configuration
    SAR_Aperture<float> ap;
    create_SARAperture<float>((float)adf, sample.data(), K, startF.data(), chirpRate.data(),
                              x.data(), y.data(), z.data(), Np, ap);
    SAR_ImageFormationParameters<float> p;
    p.update(ap);
    p.N_fft = 4096;
    p.N_x_pix = p.N_y_pix = 101;
    p.Wx_m = p.Wy_m = 10.0f;
    p.x0_m = p.y0_m = 0.0f;
    CArray<float> img((size_t)p.N_x_pix * p.N_y_pix);
    cuda_focus_SAR_image(ap, p, img);
    size_t best = 0;
    for (size_t i = 0; i < img.size(); i++)
        if (Complex<float>::abs(img[i]) > Complex<float>::abs(img[best])) best = i;
    const int bx = (int)(best / p.N_y_pix), by = (int)(best % p.N_y_pix);
    const double px = p.x0_m - p.Wx_m / 2 + bx * p.Wx_m / (p.N_x_pix - 1);
    const double py = p.y0_m - p.Wy_m / 2 + by * p.Wy_m / (p.N_y_pix - 1);
    std::printf("peak x=%.3f y=%.3f |I|=%.4f\n", px, py, Complex<float>::abs(img[best]));
    if (argc > 3) {
        std::FILE *f = std::fopen(argv[2], "wb");
        if (!f) return 3;
        std::fwrite(&img[0], sizeof(float) * 2, img.size(), f);
        std::fclose(f);
        p.dyn_range_dB = 70;
        if (writeBMPFile(p, img, argv[3]) != EXIT_SUCCESS) return 3;
    }
    return (std::fabs(px) < 1e-3 && std::fabs(py) < 1e-3) ? 0 : 2;
}
