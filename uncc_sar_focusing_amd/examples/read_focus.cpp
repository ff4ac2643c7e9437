// read_focus.cpp -- MAT-file ingest through the C++ interface, as the
// reference's command-line programs do it (src/core/charlotte_sar_api.cpp
// main path: readData -> sph_sar_data_callback_gpu -> BMP):
//
//   read_focus <in.mat> <out.bmp> [MAX_PULSES] [polarity]
//       focus on the GPU, write the BMP, print "peak x=<m> y=<m>" of the
//       default image (full unambiguous extent, N_fft = K);
//   read_focus --image <in.mat> <out.bmp> <out.bin> [MAX_PULSES] [polarity]
//       as above, and write the default image the callback forms (the same
//       create_SARAperture -> create_SARImageFormationParams ->
//       cuda_focus_SAR_image steps) with what it was formed from:
//       int32 {N_x_pix, N_y_pix, N_fft, K, Np}, float64 {x0, y0, Wx, Wy,
//       max_Wy}, float64 R0[Np], float64 startF[Np] (as the aperture holds
//       them), then the complex64 image (x-major, N_y fastest);
//   read_focus --mf <in.mat> <out.bin> <Nx> <Ny> <Wx> <Wy> [MAX_PULSES] [polarity]
//       focus_SAR_image with algorithm MATCHED_FILTER on an explicit grid
//       centred on (0, 0); writes int32 {Nx, Ny} and the complex64 image;
//   read_focus --dump <in.mat> <out.bin> [MAX_PULSES] [polarity]
//       only read: write K, Np (int32), ADF (float32), then SampleData
//       (2*K*Np float32), StartF, ChirpRate, x, y, z (Np float32 each).
#include <charlotte_sar_api.hpp>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

SAR_Aperture<float> aperture_of(PhaseHistory<float> &ph) {
    SAR_Aperture<float> ap;
    create_SARAperture<float>(ph.sph_MATData_preamble_ADF, ph.sph_MATData_Data_SampleData.data(),
                              ph.numRangeSamples, ph.sph_MATData_Data_StartF.data(),
                              ph.sph_MATData_Data_ChirpRate.data(),
                              ph.sph_MATData_Data_radarCoordinateFrame_x.data(),
                              ph.sph_MATData_Data_radarCoordinateFrame_y.data(),
                              ph.sph_MATData_Data_radarCoordinateFrame_z.data(),
                              ph.numAzimuthSamples, ap);
    return ap;
}

int callback(PhaseHistory<float> &ph) {
    return sph_sar_data_callback_gpu<float>(0, ph.sph_MATData_preamble_ADF,
                                            ph.sph_MATData_Data_SampleData.data(),
                                            ph.numRangeSamples, ph.sph_MATData_Data_StartF.data(),
                                            ph.sph_MATData_Data_ChirpRate.data(),
                                            ph.sph_MATData_Data_radarCoordinateFrame_x.data(),
                                            ph.sph_MATData_Data_radarCoordinateFrame_y.data(),
                                            ph.sph_MATData_Data_radarCoordinateFrame_z.data(),
                                            ph.numAzimuthSamples);
}

bool write_image(std::FILE *f, const CArray<float> &img) {
    return std::fwrite(&img[0], sizeof(float) * 2, img.size(), f) == img.size();
}

}  // namespace

int main(int argc, char **argv) {
    const std::string mode = argc > 1 && argv[1][0] == '-' ? argv[1] : "";
    const int a0 = mode.empty() ? 1 : 2;
    const int nfixed = mode == "--image" ? 3 : mode == "--mf" ? 6 : 2;
    if (argc < a0 + nfixed) {
        std::fprintf(stderr,
                     "usage: read_focus [--dump|--image|--mf] <in.mat> <out> [--image: <out.bin>]"
                     " [--mf: Nx Ny Wx Wy] [MAX_PULSES] [polarity]\n");
        return 1;
    }
    const std::string in = argv[a0], out = argv[a0 + 1];
    const int max_pulses = argc > a0 + nfixed ? std::atoi(argv[a0 + nfixed]) : 1 << 30;
    const std::string pol = argc > a0 + nfixed + 1 ? argv[a0 + nfixed + 1] : "any";
    PhaseHistory<float> ph;
    if (readData(in, max_pulses, pol, ph, true) != EXIT_SUCCESS) return 2;
    const int K = ph.numRangeSamples, Np = ph.numAzimuthSamples;
    if (mode == "--dump") {
        std::FILE *f = std::fopen(out.c_str(), "wb");
        if (!f) return 3;
        const int hdr[2] = {K, Np};
        std::fwrite(hdr, sizeof(int), 2, f);
        std::fwrite(&ph.sph_MATData_preamble_ADF, sizeof(float), 1, f);
        std::fwrite(ph.sph_MATData_Data_SampleData.data(), sizeof(float), 2 * (size_t)K * Np, f);
        for (const std::vector<float> *v :
             {&ph.sph_MATData_Data_StartF, &ph.sph_MATData_Data_ChirpRate,
              &ph.sph_MATData_Data_radarCoordinateFrame_x, &ph.sph_MATData_Data_radarCoordinateFrame_y,
              &ph.sph_MATData_Data_radarCoordinateFrame_z})
            std::fwrite(v->data(), sizeof(float), Np, f);
        std::fclose(f);
        return 0;
    }
    if (mode == "--mf") {
        SAR_Aperture<float> ap = aperture_of(ph);
        SAR_ImageFormationParameters<float> p;
        p.update(ap);
        p.algorithm = SAR_ImageFormationParameters<float>::MATCHED_FILTER;
        p.N_x_pix = std::atoi(argv[a0 + 2]);
        p.N_y_pix = std::atoi(argv[a0 + 3]);
        p.Wx_m = (float)std::atof(argv[a0 + 4]);
        p.Wy_m = (float)std::atof(argv[a0 + 5]);
        p.x0_m = p.y0_m = 0.0f;
        CArray<float> img((size_t)p.N_x_pix * p.N_y_pix);
        focus_SAR_image(ap, p, img);
        std::FILE *f = std::fopen(out.c_str(), "wb");
        if (!f) return 3;
        const int hdr[2] = {p.N_x_pix, p.N_y_pix};
        std::fwrite(hdr, sizeof(int), 2, f);
        const bool ok = write_image(f, img);
        std::fclose(f);
        return ok ? 0 : 3;
    }
    outputfile = out;
    if (callback(ph) != EXIT_SUCCESS) return 4;
    // the same default image, for the peak location (and the dump)
    SAR_Aperture<float> ap = aperture_of(ph);
    SAR_ImageFormationParameters<float> p;
    create_SARImageFormationParams(ap, p);
    CArray<float> img((size_t)p.N_x_pix * p.N_y_pix);
    cuda_focus_SAR_image(ap, p, img);
    if (img.size() != (size_t)p.N_x_pix * p.N_y_pix) return 5;
    size_t best = 0;
    for (size_t i = 0; i < img.size(); i++)
        if (Complex<float>::abs(img[i]) > Complex<float>::abs(img[best])) best = i;
    const int bx = (int)(best / p.N_y_pix), by = (int)(best % p.N_y_pix);
    const double px = p.x0_m - p.Wx_m / 2 + bx * (double)p.Wx_m / (p.N_x_pix - 1);
    const double py = p.y0_m - p.Wy_m / 2 + by * (double)p.Wy_m / (p.N_y_pix - 1);
    std::printf("image %dx%d dx=%.4f dy=%.4f peak x=%.4f y=%.4f\n", p.N_x_pix, p.N_y_pix,
                p.Wx_m / (p.N_x_pix - 1), p.Wy_m / (p.N_y_pix - 1), px, py);
    if (mode == "--image") {
        std::FILE *f = std::fopen(argv[a0 + 2], "wb");
        if (!f) return 3;
        const int hdr[5] = {p.N_x_pix, p.N_y_pix, p.N_fft, K, Np};
        const double geo[5] = {p.x0_m, p.y0_m, p.Wx_m, p.Wy_m, p.max_Wy_m};
        std::fwrite(hdr, sizeof(int), 5, f);
        std::fwrite(geo, sizeof(double), 5, f);
        std::vector<double> r0(Np), f0(Np);
        for (int n = 0; n < Np; n++) {
            r0[n] = ap.slant_range.data[n];
            f0[n] = ap.startF.data[n];
        }
        std::fwrite(r0.data(), sizeof(double), Np, f);
        std::fwrite(f0.data(), sizeof(double), Np, f);
        const bool ok = write_image(f, img);
        std::fclose(f);
        return ok ? 0 : 3;
    }
    return 0;
}
