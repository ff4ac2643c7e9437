// read_focus.cpp -- MAT-file ingest through the C++ interface, as the
// reference's command-line programs do it (src/core/charlotte_sar_api.cpp
// main path: readData -> sph_sar_data_callback_gpu -> BMP):
//
//   read_focus <in.mat> <out.bmp> [MAX_PULSES] [polarity]
//       focus on the GPU, write the BMP, print "peak x=<m> y=<m>" of the
//       default image (full unambiguous extent, N_fft = K);
//   read_focus --dump <in.mat> <out.bin> [MAX_PULSES] [polarity]
//       only read: write K, Np (int32), ADF (float32), then SampleData
//       (2*K*Np float32), StartF, ChirpRate, x, y, z (Np float32 each).
#include <charlotte_sar_api.hpp>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char **argv) {
    const bool dump = argc > 1 && std::string(argv[1]) == "--dump";
    const int a0 = dump ? 2 : 1;
    if (argc < a0 + 2) {
        std::fprintf(stderr, "usage: read_focus [--dump] <in.mat> <out> [MAX_PULSES] [polarity]\n");
        return 1;
    }
    const std::string in = argv[a0], out = argv[a0 + 1];
    const int max_pulses = argc > a0 + 2 ? std::atoi(argv[a0 + 2]) : 1 << 30;
    const std::string pol = argc > a0 + 3 ? argv[a0 + 3] : "any";
    PhaseHistory<float> ph;
    if (readData(in, max_pulses, pol, ph, true) != EXIT_SUCCESS) return 2;
    const int K = ph.numRangeSamples, Np = ph.numAzimuthSamples;
    if (dump) {
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
    outputfile = out;
    if (sph_sar_data_callback_gpu<float>(0, ph.sph_MATData_preamble_ADF,
                                         ph.sph_MATData_Data_SampleData.data(), K,
                                         ph.sph_MATData_Data_StartF.data(),
                                         ph.sph_MATData_Data_ChirpRate.data(),
                                         ph.sph_MATData_Data_radarCoordinateFrame_x.data(),
                                         ph.sph_MATData_Data_radarCoordinateFrame_y.data(),
                                         ph.sph_MATData_Data_radarCoordinateFrame_z.data(),
                                         Np) != EXIT_SUCCESS)
        return 4;
    // the same default image, for the peak location
    SAR_Aperture<float> ap;
    create_SARAperture<float>(ph.sph_MATData_preamble_ADF, ph.sph_MATData_Data_SampleData.data(), K,
                              ph.sph_MATData_Data_StartF.data(), ph.sph_MATData_Data_ChirpRate.data(),
                              ph.sph_MATData_Data_radarCoordinateFrame_x.data(),
                              ph.sph_MATData_Data_radarCoordinateFrame_y.data(),
                              ph.sph_MATData_Data_radarCoordinateFrame_z.data(), Np, ap);
    SAR_ImageFormationParameters<float> p;
    create_SARImageFormationParams(ap, p);
    CArray<float> img((size_t)p.N_x_pix * p.N_y_pix);
    if (sarf_focus_aperture(ap, p, img) != SARF_OK) return 5;
    size_t best = 0;
    for (size_t i = 0; i < img.size(); i++)
        if (Complex<float>::abs(img[i]) > Complex<float>::abs(img[best])) best = i;
    const int bx = (int)(best / p.N_y_pix), by = (int)(best % p.N_y_pix);
    const double px = p.x0_m - p.Wx_m / 2 + bx * (double)p.Wx_m / (p.N_x_pix - 1);
    const double py = p.y0_m - p.Wy_m / 2 + by * (double)p.Wy_m / (p.N_y_pix - 1);
    std::printf("image %dx%d dx=%.4f dy=%.4f peak x=%.4f y=%.4f\n", p.N_x_pix, p.N_y_pix,
                p.Wx_m / (p.N_x_pix - 1), p.Wy_m / (p.N_y_pix - 1), px, py);
    return 0;
}
