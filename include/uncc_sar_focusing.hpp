// uncc_sar_focusing.hpp -- the C++ data model and focusing calls of the SAR
// focusing interface, implemented as thin host code over the libsarf C ABI
// (sarf.h).  Link with -lsarf.
//
// Same names, fields and meanings as the reference
// (include/uncc_sar_focusing.hpp:34-512) so existing callers compile
// unchanged:
//   SAR_Aperture<T>                 input container (phase history K x Np
//                                   column-major, per-pulse geometry)
//   initialize_SAR_Aperture_Data    derives freq/deltaF/bandwidth/R0/azimuth
//                                   statistics (:224-393)
//   SAR_ImageFormationParameters<T> image geometry (:395-475)
//   focus_SAR_image / cuda_focus_SAR_image
//                                   range compression + backprojection, or
//                                   the matched filter (:509-512;
//                                   cpuBackProjection.hpp:230-262;
//                                   cuda_sar_focusing.hpp:526-723)
// Both focusing calls run on the MI355X through sarf_focus (backprojection)
// or sarf_focus_mf (matched filter); there is no CPU path in this library.
// Differences from the reference, all fixes: N_fft > numRangeSamples
// zero-pads the profiles (a TODO in the reference), dR and the pixel grid
// are never rounded to FP32 at kilometre ranges, the <double,double>
// instantiation works, the device is never reset.
#ifndef SARF_UNCC_SAR_FOCUSING_HPP
#define SARF_UNCC_SAR_FOCUSING_HPP

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <valarray>
#include <vector>

#include "sarf.h"
#include "uncc_complex.hpp"
#include "uncc_simplematrix.hpp"

#ifndef CLIGHT
#define CLIGHT 299792458.0 /* speed of light (m/s) */
#endif
#ifndef PI
#define PI 3.141592653589793
#endif
#ifndef CUDAFUNCTION
#define CUDAFUNCTION
#endif

template <typename T>
using Complex = unccComplex<T>;

template <typename T>
using CArray = std::valarray<Complex<T>>;

template <typename T>
class RangeBinData {
   public:
    typedef std::shared_ptr<RangeBinData> Ptr;
    SimpleMatrix<T> rangeBins;
    T minRange;
    T maxRange;
};

template <typename T>
class SAR_Aperture {
   public:
    typedef std::shared_ptr<SAR_Aperture> Ptr;

    bool format_GOTCHA;

    // GOTCHA + Sandia fields
    SimpleMatrix<Complex<T>> sampleData;  // K x Np (x polarities), column-major
    SimpleMatrix<T> Ant_x, Ant_y, Ant_z;

    // GOTCHA-only fields
    SimpleMatrix<T> freq;         // K x Np (or K) frequencies (Hz)
    SimpleMatrix<T> slant_range;  // R0 per pulse (m)
    SimpleMatrix<T> theta, phi;
    struct {
        SimpleMatrix<T> r_correct;
        SimpleMatrix<T> ph_correct;
    } af;

    // Sandia-only fields
    SimpleMatrix<T> ADF;
    SimpleMatrix<T> startF;
    SimpleMatrix<T> chirpRate;

    int polarity_channel;    // 1 HH, 2 HV, 3 VH, 4 VV
    int polarity_dimension;  // -1: a single polarity

    // derived by initialize_SAR_Aperture_Data
    int numRangeSamples = 0;
    int numAzimuthSamples = 0;
    int numPolarities = 0;
    SimpleMatrix<T> bandwidth;
    SimpleMatrix<T> deltaF;
    T mean_startF = 0, mean_deltaF = 0, mean_bandwidth = 0;
    SimpleMatrix<T> Ant_Az, Ant_deltaAz, Ant_El, Ant_deltaEl;
    T mean_Ant_El = 0, mean_deltaAz = 0, mean_deltaEl = 0, mean_Ant_deltaAz = 0,
      mean_Ant_deltaEl = 0, Ant_totalAz = 0, Ant_totalEl = 0;

    SAR_Aperture() : format_GOTCHA(true), polarity_channel(1), polarity_dimension(-1) {}
    virtual ~SAR_Aperture() {}

    Ptr create() { return std::make_shared<SAR_Aperture>(); }

    // Copies one polarity channel (and the derived fields) into dst.
    int exportData(SAR_Aperture<T> &dst, int channel) {
        dst.numRangeSamples = numRangeSamples;
        dst.numAzimuthSamples = numAzimuthSamples;
        dst.numPolarities = 1;
        dst.sampleData = sampleData.getRange(
            {Range(0, numRangeSamples), Range(0, numAzimuthSamples), Range(channel, channel + 1)});
        const std::vector<Range> az = {Range(0, 1), Range(0, numAzimuthSamples),
                                       Range(channel, channel + 1)};
        dst.Ant_x = Ant_x.getRange(az);
        dst.Ant_y = Ant_y.getRange(az);
        dst.Ant_z = Ant_z.getRange(az);
        dst.startF = startF.getRange(az);
        dst.chirpRate = chirpRate.getRange(az);
        dst.bandwidth = bandwidth;
        dst.freq = freq;
        dst.slant_range = slant_range;
        dst.deltaF = deltaF;
        dst.Ant_Az = Ant_Az;
        dst.Ant_El = Ant_El;
        dst.Ant_deltaAz = Ant_deltaAz;
        dst.Ant_deltaEl = Ant_deltaEl;
        dst.mean_Ant_El = mean_Ant_El;
        dst.mean_Ant_deltaAz = mean_Ant_deltaAz;
        dst.mean_deltaAz = mean_deltaAz;
        dst.mean_deltaEl = mean_deltaEl;
        dst.mean_startF = mean_startF;
        dst.mean_deltaF = mean_deltaF;
        dst.mean_bandwidth = mean_bandwidth;
        dst.Ant_totalAz = Ant_totalAz;
        dst.Ant_totalEl = Ant_totalEl;
        return EXIT_SUCCESS;
    }
};

template <typename T>
inline std::ostream &operator<<(std::ostream &os, const SAR_Aperture<T> &c) {
    os << "sampleData" << c.sampleData << "\nAnt_x" << c.Ant_x << "\nAnt_y" << c.Ant_y
       << "\nAnt_z" << c.Ant_z << "\nfreq" << c.freq << "\nStartF" << c.startF << "\ndeltaF"
       << c.deltaF << "\nbandwidth" << c.bandwidth << "\nslant_range" << c.slant_range
       << "\nAnt_totalAz = " << c.Ant_totalAz << "\nAnt_totalEl = " << c.Ant_totalEl << "\n";
    return os;
}

namespace sarf_detail {
template <typename M>
inline int per_polarity(const M &m, int pol_dim) {
    const int n = m.nelem();
    if (pol_dim >= 0 && pol_dim < (int)m.shape.size() && m.shape[pol_dim] > 0)
        return n / m.shape[pol_dim];
    return n;
}
template <typename T>
inline T mean_of(const std::vector<T> &v) {
    if (v.empty()) return T(0);
    double s = 0.0;
    for (const T &x : v) s += (double)x;
    return (T)(s / (double)v.size());
}
// Image width of the default image, in the reference's operation order
// (uncc_sar_focusing.hpp:461, charlotte_sar_api.hpp:189): the float product
// Wx * N_y is truncated to int first, then divided by Wy in the precision
// of T, and the quotient truncated again when it is stored.  Dividing
// before truncating gives a different width for about 0.3 % of geometries.
template <typename T>
inline int square_pixel_columns(T Wx, int Ny, T Wy) {
    const int wx_ny = (int)((float)Wx * Ny);
    const T cols = wx_ny / Wy;
    return (int)cols;
}
}  // namespace sarf_detail

// Fills the derived fields of an aperture (uncc_sar_focusing.hpp:224-393):
// per-pulse frequencies (from one frequency vector, or from start frequency,
// chirp rate and A/D rate), R0 = |a_n| when absent, deltaF, bandwidth, their
// means, and the azimuth/elevation statistics of the phase centres.
template <typename T>
int initialize_SAR_Aperture_Data(SAR_Aperture<T> &ap) {
    using sarf_detail::per_polarity;
    if (ap.sampleData.shape.size() < 2) {
        std::cerr << "initialize_SAR_Aperture_Data: sampleData must be K x Np" << std::endl;
        return EXIT_FAILURE;
    }
    const int K = ap.sampleData.shape[0], Np = ap.sampleData.shape[1];
    ap.numRangeSamples = K;
    ap.numAzimuthSamples = Np;
    ap.numPolarities = ap.sampleData.shape.size() > 2 ? ap.sampleData.shape[2] : 1;
    const int pd = ap.polarity_dimension;
    if (per_polarity(ap.Ant_x, pd) != Np || per_polarity(ap.Ant_y, pd) != Np ||
        per_polarity(ap.Ant_z, pd) != Np) {
        std::cerr << "initialize_SAR_Aperture_Data: need one antenna position per pulse"
                  << std::endl;
        return EXIT_FAILURE;
    }
    if (per_polarity(ap.freq, pd) != K * Np) {
        if (!ap.freq.isEmpty() && ap.freq.shape[0] == K && K > 1) {
            // one frequency vector for every pulse
            std::vector<T> f1(ap.freq.data.begin(), ap.freq.data.begin() + K);
            const T fmin = *std::min_element(f1.begin(), f1.end());
            const T fmax = *std::max_element(f1.begin(), f1.end());
            const T df = std::abs(f1[1] - f1[0]);
            const bool fill_start = per_polarity(ap.startF, pd) != Np;
            const bool fill_df = per_polarity(ap.deltaF, pd) != Np;
            ap.freq.shape = {K, Np};
            ap.freq.data.clear();
            if (fill_start) ap.startF = SimpleMatrix<T>{{Np}, {}};
            if (fill_df) ap.deltaF = SimpleMatrix<T>{{Np}, {}};
            ap.bandwidth = SimpleMatrix<T>{{Np}, {}};
            for (int n = 0; n < Np; n++) {
                ap.freq.data.insert(ap.freq.data.end(), f1.begin(), f1.end());
                ap.bandwidth.data.push_back(fmax - fmin);
                if (fill_start) ap.startF.data.push_back(fmin);
                if (fill_df) ap.deltaF.data.push_back(df);
            }
        } else if (!ap.startF.isEmpty() && ap.startF.nelem() == Np && !ap.ADF.isEmpty() &&
                   ap.ADF.nelem() == 1 && !ap.chirpRate.isEmpty() && ap.chirpRate.nelem() == Np) {
            // Sandia: f[n,k] = startF[n] + k chirpRate[n] / ADF
            ap.freq = SimpleMatrix<T>{{K, Np}, {}};
            ap.deltaF = SimpleMatrix<T>{{Np}, {}};
            ap.bandwidth = SimpleMatrix<T>{{Np}, {}};
            for (int n = 0; n < Np; n++) {
                const T df = ap.chirpRate.data[n] / ap.ADF.data[0];
                for (int k = 0; k < K; k++) ap.freq.data.push_back(ap.startF.data[n] + k * df);
                ap.deltaF.data.push_back(df);
                ap.bandwidth.data.push_back((K - 1) * df);
            }
        }
    }
    if (per_polarity(ap.startF, pd) != Np && per_polarity(ap.freq, pd) == K * Np) {
        ap.startF = SimpleMatrix<T>{{Np}, {}};
        for (int n = 0; n < Np; n++) {
            const auto b = ap.freq.data.begin() + (size_t)n * K;
            ap.startF.data.push_back(*std::min_element(b, b + K));
        }
    }
    if (per_polarity(ap.slant_range, pd) != Np) {
        ap.slant_range = SimpleMatrix<T>{{Np}, {}};
        for (int n = 0; n < Np; n++) {
            const double x = ap.Ant_x.data[n], y = ap.Ant_y.data[n], z = ap.Ant_z.data[n];
            ap.slant_range.data.push_back((T)std::sqrt(x * x + y * y + z * z));
        }
    }
    if (per_polarity(ap.deltaF, pd) != Np && per_polarity(ap.freq, pd) == K * Np && K > 1) {
        ap.deltaF = SimpleMatrix<T>{{Np}, {}};
        for (int n = 0; n < Np; n++)
            ap.deltaF.data.push_back(
                std::abs(ap.freq.data[(size_t)n * K + 1] - ap.freq.data[(size_t)n * K]));
    }
    if (ap.bandwidth.nelem() != Np && !ap.deltaF.data.empty()) {
        ap.bandwidth = SimpleMatrix<T>{{Np}, {}};
        for (int n = 0; n < Np; n++) ap.bandwidth.data.push_back((K - 1) * ap.deltaF.data[n]);
    }
    ap.mean_startF = sarf_detail::mean_of(ap.startF.data);
    ap.mean_deltaF = sarf_detail::mean_of(ap.deltaF.data);
    ap.mean_bandwidth = sarf_detail::mean_of(ap.bandwidth.data);
    if (per_polarity(ap.Ant_Az, pd) != Np) {
        ap.Ant_Az = SimpleMatrix<T>{{Np}, {}};
        ap.Ant_El = SimpleMatrix<T>{{Np}, {}};
        ap.Ant_deltaAz = SimpleMatrix<T>{{std::max(Np - 1, 0)}, {}};
        ap.Ant_deltaEl = SimpleMatrix<T>{{std::max(Np - 1, 0)}, {}};
        for (int n = 0; n < Np; n++) {
            const double x = ap.Ant_x.data[n], y = ap.Ant_y.data[n], z = ap.Ant_z.data[n];
            ap.Ant_Az.data.push_back((T)std::atan2(y, x));
            ap.Ant_El.data.push_back((T)std::atan2(z, std::sqrt(x * x + y * y)));
            if (n > 0) {
                ap.Ant_deltaAz.data.push_back(ap.Ant_Az.data[n] - ap.Ant_Az.data[n - 1]);
                ap.Ant_deltaEl.data.push_back(ap.Ant_El.data[n] - ap.Ant_El.data[n - 1]);
            }
        }
        ap.mean_Ant_deltaAz = sarf_detail::mean_of(ap.Ant_deltaAz.data);
        ap.mean_Ant_El = sarf_detail::mean_of(ap.Ant_El.data);
        ap.mean_Ant_deltaEl = sarf_detail::mean_of(ap.Ant_deltaEl.data);
        const auto az = std::minmax_element(ap.Ant_Az.data.begin(), ap.Ant_Az.data.end());
        const auto el = std::minmax_element(ap.Ant_El.data.begin(), ap.Ant_El.data.end());
        ap.Ant_totalAz = *az.second - *az.first;
        ap.Ant_totalEl = *el.second - *el.first;
    }
    return EXIT_SUCCESS;
}

template <typename T>
class SAR_ImageFormationParameters {
   public:
    enum ALGORITHM { UNKNOWN, MATCHED_FILTER, BACKPROJECTION, PHASE_GRADIENT_ALGORITHM } algorithm;
    int pol = 0;
    int N_fft;          // range-profile length (>= numRangeSamples: zero padding)
    bool zeropad_fft;
    int N_x_pix, N_y_pix;
    T x0_m, y0_m;       // image centre (m)
    T Wx_m = 0, Wy_m = 0;          // image extent (m); pixel spacing W/(N-1)
    T max_Wx_m = 0, max_Wy_m = 0;  // unambiguous extents; max_Wy_m = c / (2 mean deltaF)
    T dyn_range_dB;
    T slant_rangeResolution = 0, ground_rangeResolution = 0, azimuthResolution = 0;

    SAR_ImageFormationParameters()
        : algorithm(BACKPROJECTION), N_fft(512), zeropad_fft(true), N_x_pix(512), N_y_pix(512),
          x0_m(0), y0_m(0), dyn_range_dB(70) {}

    template <typename A>
    void update(const SAR_Aperture<A> &ap) {
        max_Wy_m = (T)(CLIGHT / (2.0 * ap.mean_deltaF));
        max_Wx_m = (T)(CLIGHT / (2.0 * std::abs(ap.mean_Ant_deltaAz) * ap.mean_startF));
        slant_rangeResolution = (T)(CLIGHT / (2.0 * ap.mean_bandwidth));
        ground_rangeResolution = (T)(slant_rangeResolution / std::sin(ap.mean_Ant_El));
        azimuthResolution = (T)(CLIGHT / (2.0 * ap.Ant_totalAz * ap.mean_startF));
    }

    template <typename MyT, typename A>
    static SAR_ImageFormationParameters<MyT> create(const SAR_Aperture<A> &ap) {
        SAR_ImageFormationParameters<MyT> p;
        p.N_fft = ap.numRangeSamples;
        p.N_y_pix = p.N_fft;
        p.update(ap);
        p.Wx_m = p.max_Wx_m;
        p.Wy_m = p.max_Wy_m;
        p.N_x_pix = sarf_detail::square_pixel_columns<MyT>(p.Wx_m, p.N_y_pix, p.Wy_m);
        return p;
    }
};

template <typename T>
inline std::ostream &operator<<(std::ostream &os, const SAR_ImageFormationParameters<T> &c) {
    os << "Nfft = {" << c.N_fft << "}\nN_x_pix = {" << c.N_x_pix << "}\nN_y_pix = {" << c.N_y_pix
       << "}\nx0_m = {" << c.x0_m << "}\ny0_m = {" << c.y0_m << "}\nmax_Wx_m = {" << c.max_Wx_m
       << "}\nmax_Wy_m = {" << c.max_Wy_m << "}\nWx_m = {" << c.Wx_m << "}\nWy_m = {" << c.Wy_m
       << "}\n";
    return os;
}

// Process-wide libsarf context on HIP device 0 (created on first use).
inline sarf_ctx *sarf_default_context() {
    static std::once_flag once;
    static sarf_ctx *ctx = nullptr;
    std::call_once(once, [] {
        if (sarf_ctx_create(0, &ctx) != SARF_OK) {
            std::cerr << "libsarf: " << sarf_last_error() << std::endl;
            ctx = nullptr;
        }
    });
    return ctx;
}

// Range compression + backprojection of `sar_data` onto the image of
// `params` (output: N_y_pix x N_x_pix complex, y fastest), on the GPU.
// Returns SARF_OK or a negative sarf.h code (message in sarf_last_error()).
template <typename T, typename P>
int sarf_focus_aperture(const SAR_Aperture<T> &ap, const SAR_ImageFormationParameters<P> &p,
                        CArray<T> &out, int precision = SARF_BP_MIXED) {
    sarf_ctx *ctx = sarf_default_context();
    if (!ctx) return SARF_ERR_NODEVICE;
    const int K = ap.numRangeSamples, Np = ap.numAzimuthSamples;
    if (K <= 0 || Np <= 0 || (int)ap.sampleData.data.size() < K * Np ||
        (int)ap.Ant_x.data.size() < Np || (int)ap.slant_range.data.size() < Np ||
        (int)ap.startF.data.size() < Np)
        return SARF_ERR_INVALID;
    std::vector<float> ph(2 * (size_t)K * Np);
    for (size_t i = 0; i < (size_t)K * Np; i++) {
        ph[2 * i] = (float)ap.sampleData.data[i]._M_real;
        ph[2 * i + 1] = (float)ap.sampleData.data[i]._M_imag;
    }
    std::vector<double> ant(3 * (size_t)Np), r0(Np), f0(Np);
    for (int n = 0; n < Np; n++) {
        ant[3 * n] = ap.Ant_x.data[n];
        ant[3 * n + 1] = ap.Ant_y.data[n];
        ant[3 * n + 2] = ap.Ant_z.data[n];
        r0[n] = ap.slant_range.data[n];
        f0[n] = ap.startF.data[n];
    }
    sarf_focus_desc d{};
    d.K = K;
    d.Np = Np;
    d.Nfft = std::max(p.N_fft, K);
    d.Nx = p.N_x_pix;
    d.Ny = p.N_y_pix;
    d.x0 = p.x0_m;
    d.y0 = p.y0_m;
    d.Wx = p.Wx_m;
    d.Wy = p.Wy_m;
    d.maxWr = p.max_Wy_m;
    d.precision = precision;
    std::vector<float> img(2 * (size_t)d.Nx * d.Ny);
    const int rc = sarf_focus(ctx, &d, ph.data(), ant.data(), r0.data(), f0.data(), img.data());
    if (rc != SARF_OK) return rc;
    if (out.size() != (size_t)d.Nx * d.Ny) out.resize((size_t)d.Nx * d.Ny);
    for (size_t i = 0; i < out.size(); i++) out[i] = Complex<T>((T)img[2 * i], (T)img[2 * i + 1]);
    return SARF_OK;
}

// Matched-filter image of `ap` on the pixel grid of `p` (output: N_y_pix x
// N_x_pix complex, y fastest), on the GPU: what run_mf computes
// (cpuBackProjection.hpp:165-228) -- every pixel, every sample, no range
// compression and no swath gate -- over the pulses run_mf visits, index 2
// to Np-1 (:218).  The frequencies of pulse n are the stepped ramp through
// freq[n,0] and freq[n,K-1] (both record formats are stepped-frequency:
// uncc_sar_focusing.hpp:246-308).  Returns SARF_OK or a negative sarf.h code.
template <typename T, typename P>
int sarf_mf_aperture(const SAR_Aperture<T> &ap, const SAR_ImageFormationParameters<P> &p,
                     CArray<T> &out) {
    sarf_ctx *ctx = sarf_default_context();
    if (!ctx) return SARF_ERR_NODEVICE;
    const int K = ap.numRangeSamples, Np = ap.numAzimuthSamples;
    if (K <= 0 || Np <= 0 || (int)ap.sampleData.data.size() < K * Np ||
        (int)ap.Ant_x.data.size() < Np || (int)ap.slant_range.data.size() < Np ||
        (int)ap.freq.data.size() < K * Np || p.N_x_pix < 2 || p.N_y_pix < 2)
        return SARF_ERR_INVALID;
    const size_t npix = (size_t)p.N_x_pix * p.N_y_pix;
    if (out.size() != npix) out.resize(npix);
    const int n0 = 2;  // run_mf skips pulses 0 and 1
    const int np = Np - n0;
    if (np <= 0) {
        for (size_t i = 0; i < npix; i++) out[i] = Complex<T>(T(0), T(0));
        return SARF_OK;
    }
    std::vector<float> ph(2 * (size_t)K * np);
    for (size_t i = 0; i < (size_t)K * np; i++) {
        const auto &v = ap.sampleData.data[(size_t)n0 * K + i];
        ph[2 * i] = (float)v._M_real;
        ph[2 * i + 1] = (float)v._M_imag;
    }
    std::vector<double> ant(3 * (size_t)np), r0(np), f0(np), df(np);
    for (int n = 0; n < np; n++) {
        const int m = n0 + n;
        ant[3 * n] = ap.Ant_x.data[m];
        ant[3 * n + 1] = ap.Ant_y.data[m];
        ant[3 * n + 2] = ap.Ant_z.data[m];
        r0[n] = ap.slant_range.data[m];
        f0[n] = ap.freq.data[(size_t)m * K];
        df[n] = K > 1 ? ((double)ap.freq.data[(size_t)m * K + K - 1] - f0[n]) / (K - 1) : 0.0;
    }
    sarf_mf_desc d{};
    d.Nx = p.N_x_pix;
    d.Ny = p.N_y_pix;
    d.K = K;
    d.Np = np;
    d.x0 = p.x0_m;
    d.y0 = p.y0_m;
    d.Wx = p.Wx_m;
    d.Wy = p.Wy_m;
    std::vector<float> img(2 * npix);
    const int rc = sarf_focus_mf(ctx, &d, ph.data(), ant.data(), r0.data(), f0.data(), df.data(),
                                 img.data());
    if (rc != SARF_OK) return rc;
    for (size_t i = 0; i < npix; i++) out[i] = Complex<T>((T)img[2 * i], (T)img[2 * i + 1]);
    return SARF_OK;
}

// Focusing call of the GPU build (cuda_sar_focusing.hpp:526-723): range
// compression + backprojection, or (which the reference's GPU entry refuses,
// :537-545) the matched filter of run_mf.
template <typename T, typename P>
void cuda_focus_SAR_image(const SAR_Aperture<T> &sar_data,
                          const SAR_ImageFormationParameters<P> &sar_image_params,
                          CArray<T> &output_image) {
    int rc;
    switch (sar_image_params.algorithm) {
        case SAR_ImageFormationParameters<P>::BACKPROJECTION:
            rc = sarf_focus_aperture(sar_data, sar_image_params, output_image);
            break;
        case SAR_ImageFormationParameters<P>::MATCHED_FILTER:
            rc = sarf_mf_aperture(sar_data, sar_image_params, output_image);
            break;
        default:
            std::cerr << "focus_SAR_image: algorithm requested is not available" << std::endl;
            return;
    }
    if (rc != SARF_OK)
        std::cerr << "focus_SAR_image: " << (rc == SARF_ERR_INVALID && !*sarf_last_error()
                                                 ? "aperture or image parameters incomplete"
                                                 : sarf_last_error())
                  << std::endl;
}

// Focusing entry of the reference's CPU build (uncc_sar_focusing.hpp:509-512,
// cpuBackProjection.hpp:230-262): BACKPROJECTION -> run_bp, MATCHED_FILTER ->
// run_mf (:249-257); here both run on the GPU.
template <typename T, typename P>
void focus_SAR_image(const SAR_Aperture<T> &SARData, const SAR_ImageFormationParameters<P> &SARImgParams,
                     CArray<T> &output_image) {
    cuda_focus_SAR_image(SARData, SARImgParams, output_image);
}

#endif  // SARF_UNCC_SAR_FOCUSING_HPP
