// charlotte_sar_api.hpp -- the callback API of the SAR focusing interface
// (reference include/charlotte_sar_api.hpp:35-316), as thin host code over
// the libsarf C ABI.  Link with -lsarf.
//
//   PhaseHistory<T>                Sandia phase-history record
//   sph_sar_data_callback_{cpu,gpu}<T>(id, ADF, SampleData, K, StartF,
//       ChirpRate, x, y, z, Np)   raw arrays -> aperture -> image -> BMP file
//   sar_data_callback<T>           both of the above
//   create_SARAperture, create_SARImageFormationParams, writeBMPFile
//
// SampleData is interleaved re/im, K samples per pulse, pulses contiguous.
// Both callbacks focus on the GPU (this library has no CPU path).  The
// globals `outputfile` and `verbose` are defined in libsarf.
#ifndef SARF_CHARLOTTE_SAR_API_HPP
#define SARF_CHARLOTTE_SAR_API_HPP

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "uncc_sar_focusing.hpp"

extern std::string outputfile;
extern bool verbose;

template <typename T>
class PhaseHistory {
   public:
    int id;
    T sph_MATData_preamble_ADF;
    std::vector<T> sph_MATData_Data_SampleData;
    int numRangeSamples;
    std::vector<T> sph_MATData_Data_StartF;
    std::vector<T> sph_MATData_Data_ChirpRate;
    std::vector<T> sph_MATData_Data_radarCoordinateFrame_x;
    std::vector<T> sph_MATData_Data_radarCoordinateFrame_y;
    std::vector<T> sph_MATData_Data_radarCoordinateFrame_z;
    int numAzimuthSamples;
};

// MAT-file input (charlotte_sar_api.cpp:36-148): reads a Level 5 MAT-file
// (Sandia sph_MATData.* or GOTCHA data.* record) with libsarf's own reader
// and fills `ph` with the central min(MAX_PULSES, Np) pulses of the channel
// named by `polarity` ("HH"/"any", "HV", "VH", "VV").
int readData(const std::string &inputfile, const int MAX_PULSES, const std::string &polarity,
             PhaseHistory<float> &ph, bool verbose);

using NumericType = float;
using ComplexType = Complex<NumericType>;
using ComplexArrayType = CArray<NumericType>;

// Display image and file in the reference's format (uncc_sar_globals.hpp:
// 66-118 with loadbmp.h:182-245): |I| in dB over [-dyn_range_dB, 0] mapped to
// 0..255 in float, far range (largest y) in the top row, written as a
// 24-bit grey BMP, rows bottom-up and padded to 4 bytes, header size field
// 54 + 3*W*H.
template <typename T, typename P>
int writeBMPFile(const SAR_ImageFormationParameters<P> &params, const CArray<T> &img,
                 const std::string &filename) {
    const int W = params.N_x_pix, H = params.N_y_pix;
    if ((long)img.size() < (long)W * H || W <= 0 || H <= 0) return EXIT_FAILURE;
    float vmax = 0.0f;
    for (size_t i = 0; i < img.size(); i++) vmax = std::max(vmax, (float)Complex<T>::abs(img[i]));
    const float dyn = (float)params.dyn_range_dB;
    const float scale = (float)(255.0 / dyn);
    // top-down display rows: row r shows y = H-1-r
    std::vector<unsigned char> rows((size_t)W * H);
    for (int r = 0; r < H; r++) {
        const int y = H - 1 - r;
        for (int x = 0; x < W; x++) {
            const float a = (float)Complex<T>::abs(img[(size_t)x * H + y]);
            const float l = std::log10(a / vmax);
            volatile float t = 20.0f * l;  // no contraction: same rounding as the reference
            volatile float u = t + dyn;
            const float v = scale * u;
            rows[(size_t)r * W + x] = (v < 0.0f || !(v == v)) ? 0 : (unsigned char)v;
        }
    }
    const uint32_t size = 54u + 3u * (uint32_t)W * (uint32_t)H;
    unsigned char hdr[54] = {'B', 'M'};
    auto put32 = [&](int at, uint32_t v) {
        for (int b = 0; b < 4; b++) hdr[at + b] = (unsigned char)(v >> (8 * b));
    };
    put32(2, size);
    hdr[10] = 54;
    hdr[14] = 40;
    put32(18, (uint32_t)W);
    put32(22, (uint32_t)H);
    hdr[26] = 1;
    hdr[28] = 24;
    const int pad = (4 - (W * 3) % 4) % 4;
    std::vector<unsigned char> body;
    body.reserve((size_t)(3 * W + pad) * H);
    for (int r = H - 1; r >= 0; r--) {  // bottom-up
        for (int x = 0; x < W; x++) body.insert(body.end(), 3, rows[(size_t)r * W + x]);
        body.insert(body.end(), pad, 0);
    }
    std::ofstream f(filename, std::ios::binary);
    if (!f) return EXIT_FAILURE;
    f.write((const char *)hdr, sizeof hdr);
    f.write((const char *)body.data(), body.size());
    return f ? EXIT_SUCCESS : EXIT_FAILURE;
}

template <typename T>
int create_SARAperture(T ADF, T *SampleData, int numRangeSamples, T *StartF, T *ChirpRate, T *x,
                       T *y, T *z, int numAzimuthSamples, SAR_Aperture<T> &ap) {
    ap.ADF.shape.push_back(1);
    ap.ADF.data.push_back(ADF);
    int np[1] = {numAzimuthSamples};
    int dims[2] = {numRangeSamples, numAzimuthSamples};
    import_MatrixComplex<T, Complex<T>>(SampleData, dims, 2, ap.sampleData);
    import_Vector<T, T>(StartF, np, 1, ap.startF);
    import_Vector<T, T>(ChirpRate, np, 1, ap.chirpRate);
    import_Vector<T, T>(x, np, 1, ap.Ant_x);
    import_Vector<T, T>(y, np, 1, ap.Ant_y);
    import_Vector<T, T>(z, np, 1, ap.Ant_z);
    ap.format_GOTCHA = false;
    return initialize_SAR_Aperture_Data(ap);
}

// Default image: N_fft = K (no padding), N_y = K, full unambiguous extent,
// N_x scaled for square pixels (charlotte_sar_api.hpp:168-197).
template <typename T>
int create_SARImageFormationParams(const SAR_Aperture<T> &ap, SAR_ImageFormationParameters<T> &p) {
    p.N_fft = ap.numRangeSamples;
    p.N_x_pix = ap.numAzimuthSamples;
    p.N_y_pix = ap.numRangeSamples;
    p.update(ap);
    p.Wx_m = p.max_Wx_m;
    p.Wy_m = p.max_Wy_m;
    p.N_x_pix = sarf_detail::square_pixel_columns<T>(p.Wx_m, p.N_y_pix, p.Wy_m);
    return EXIT_SUCCESS;
}

template <typename T>
int sph_sar_data_callback_gpu(int id, T ADF, T *SampleData, int numRangeSamples, T *StartF,
                              T *ChirpRate, T *x, T *y, T *z, int numAzimuthSamples) {
    (void)id;
    SAR_Aperture<T> ap;
    if (create_SARAperture(ADF, SampleData, numRangeSamples, StartF, ChirpRate, x, y, z,
                           numAzimuthSamples, ap) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    if (verbose) std::cout << "Data for focusing:\n" << ap << std::endl;
    SAR_ImageFormationParameters<T> p;
    create_SARImageFormationParams(ap, p);
    if (verbose) std::cout << p << std::endl;
    CArray<T> img((size_t)p.N_y_pix * p.N_x_pix);
    const int rc = sarf_focus_aperture(ap, p, img);
    if (rc != SARF_OK) {
        std::cerr << "sph_sar_data_callback_gpu: " << sarf_last_error() << std::endl;
        return EXIT_FAILURE;
    }
    p.dyn_range_dB = 70;
    return writeBMPFile(p, img, outputfile);
}

// The reference's CPU callback; this library forms the image on the GPU.
template <typename T>
int sph_sar_data_callback_cpu(int id, T ADF, T *SampleData, int numRangeSamples, T *StartF,
                              T *ChirpRate, T *x, T *y, T *z, int numAzimuthSamples) {
    return sph_sar_data_callback_gpu(id, ADF, SampleData, numRangeSamples, StartF, ChirpRate, x, y,
                                     z, numAzimuthSamples);
}

template <typename T>
int sar_data_callback(int id, T ADF, T *SampleData, int numRangeSamples, T *StartF, T *ChirpRate,
                      T *x, T *y, T *z, int numAzimuthSamples) {
    sph_sar_data_callback_cpu<T>(id, ADF, SampleData, numRangeSamples, StartF, ChirpRate, x, y, z,
                                 numAzimuthSamples);
    sph_sar_data_callback_gpu<T>(id, ADF, SampleData, numRangeSamples, StartF, ChirpRate, x, y, z,
                                 numAzimuthSamples);
    return EXIT_SUCCESS;
}

#endif  // SARF_CHARLOTTE_SAR_API_HPP
