// Host check of the default-image width of the C++ interface headers.
//
//   nxpix_check aperture <in.bin> <out.bin>
//     in:  records of 4 float32 {K, mean_deltaF, mean_Ant_deltaAz, mean_startF}
//     out: per record, from create_SARImageFormationParams (charlotte_sar_api.hpp)
//          and SAR_ImageFormationParameters<float>::create / <double>::create
//          (uncc_sar_focusing.hpp): float32 {Wx, Wy}, int32 N_x_pix, then
//          float32 {Wx, Wy}, int32 N_x_pix, then float64 {Wx, Wy}, int32 N_x_pix
//   nxpix_check helper <in.bin> <out.bin>
//     in:  records of {float32 Wx, int32 Ny, float32 Wy}
//     out: int32 N_x for T = float, int32 N_x for T = double (Wx, Wy widened)
#include <charlotte_sar_api.hpp>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

std::string outputfile;
bool verbose = false;

static std::vector<char> slurp(const char *path) {
    std::vector<char> buf;
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return buf;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    return buf;
}

template <typename V>
static void put(std::vector<char> &out, V v) {
    const char *p = (const char *)&v;
    out.insert(out.end(), p, p + sizeof v);
}

int main(int argc, char **argv) {
    if (argc != 4) return 1;
    const std::string mode = argv[1];
    const std::vector<char> in = slurp(argv[2]);
    std::vector<char> out;
    if (mode == "aperture") {
        const size_t n = in.size() / (4 * sizeof(float));
        const float *r = (const float *)in.data();
        for (size_t i = 0; i < n; i++, r += 4) {
            SAR_Aperture<float> ap;
            ap.numRangeSamples = (int)r[0];
            ap.numAzimuthSamples = 128;
            ap.mean_deltaF = r[1];
            ap.mean_Ant_deltaAz = r[2];
            ap.mean_startF = r[3];
            ap.mean_bandwidth = r[1] * (r[0] - 1);
            ap.mean_Ant_El = 0.6f;
            ap.Ant_totalAz = r[2] * 127;
            SAR_ImageFormationParameters<float> a;
            create_SARImageFormationParams(ap, a);
            put(out, (float)a.Wx_m);
            put(out, (float)a.Wy_m);
            put(out, (int)a.N_x_pix);
            const auto b = SAR_ImageFormationParameters<float>::create<float>(ap);
            put(out, (float)b.Wx_m);
            put(out, (float)b.Wy_m);
            put(out, (int)b.N_x_pix);
            const auto c = SAR_ImageFormationParameters<double>::create<double>(ap);
            put(out, (double)c.Wx_m);
            put(out, (double)c.Wy_m);
            put(out, (int)c.N_x_pix);
        }
    } else if (mode == "helper") {
        const size_t rec = 2 * sizeof(float) + sizeof(int);
        for (size_t i = 0; i + rec <= in.size(); i += rec) {
            float wx, wy;
            int ny;
            std::memcpy(&wx, in.data() + i, 4);
            std::memcpy(&ny, in.data() + i + 4, 4);
            std::memcpy(&wy, in.data() + i + 8, 4);
            put(out, sarf_detail::square_pixel_columns<float>(wx, ny, wy));
            put(out, sarf_detail::square_pixel_columns<double>(wx, ny, wy));
        }
    } else {
        return 1;
    }
    std::FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 2;
    std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    std::printf("nxpix_check ok %zu bytes\n", out.size());
    return 0;
}
