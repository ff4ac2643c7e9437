// sarf_cxx_api.cpp -- definitions the C++ callback interface
// (include/charlotte_sar_api.hpp) needs from the library: the globals and
// readData, the MAT-file ingest of src/core/charlotte_sar_api.cpp:36-148.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <map>
#include <string>

#include "charlotte_sar_api.hpp"
#include "sarf_matv5.hpp"

// Output file of the sph_sar_data_callback_* functions and the verbosity
// flag (charlotte_sar_api.hpp:32-33 of the reference).
std::string outputfile = "sar_image.bmp";
bool verbose = false;

namespace {

using sarf::matv5::Array;

const Array *find(const std::map<std::string, Array> &v, const std::string &k) {
    auto it = v.find(k);
    return it == v.end() || it->second.numel() == 0 ? nullptr : &it->second;
}

// Per-pulse value n of a vector that holds Np values, or Np values per
// polarity channel (Sandia files may store one set per channel).
bool per_pulse(const Array *a, int Np, int npol, int ch, std::vector<double> &out) {
    if (!a) return false;
    const size_t n = a->numel();
    size_t off = 0;
    if (n == (size_t)Np * npol && npol > 1) off = (size_t)ch * Np;
    else if (n < (size_t)Np) return false;
    out.assign(a->re.begin() + off, a->re.begin() + off + Np);
    return true;
}

}  // namespace

// Reads a Sandia phase-history record (sph_MATData.preamble.ADF,
// sph_MATData.Data.{SampleData, StartF, ChirpRate, radarCoordinateFrame.x/y/z};
// uncc_sar_matio.cpp:40-52) or a GOTCHA record (data.{fp, freq, x, y, z};
// :23-34) from a Level 5 MAT-file and fills `ph` with the central
// min(MAX_PULSES, Np) pulses (charlotte_sar_api.cpp:86-95) of polarity
// channel HH/any = 0, HV = 1, VH = 2, VV = 3 (the channel selected by
// `polarity`; the reference exports channel 2 for every Sandia request,
// :79-80).  GOTCHA frequencies become StartF = min(freq) per pulse and
// ChirpRate = deltaF with ADF = 1, so the callback recovers them.
int readData(const std::string &inputfile, const int MAX_PULSES, const std::string &polarity,
             PhaseHistory<float> &ph, bool verbose_) {
    std::map<std::string, Array> v;
    std::string err;
    if (sarf::matv5::read_file(inputfile, v, err) != 0) {
        std::cerr << "readData: " << err << std::endl;
        return EXIT_FAILURE;
    }
    const Array *sd = find(v, "sph_MATData.Data.SampleData");
    const bool sandia = sd != nullptr;
    if (!sd) sd = find(v, "data.fp");
    if (!sd || !sd->complex() || sd->dims.size() < 2) {
        std::cerr << "readData: " << inputfile
                  << " holds neither sph_MATData.Data.SampleData nor data.fp (complex)" << std::endl;
        return EXIT_FAILURE;
    }
    const int K = sd->dims[0], Np = sd->dims[1];
    const int npol = sd->dims.size() > 2 ? sd->dims[2] : 1;
    int ch = -1;
    if (polarity == "HH" || polarity == "any") ch = 0;
    else if (polarity == "HV") ch = 1;
    else if (polarity == "VH") ch = 2;
    else if (polarity == "VV") ch = 3;
    if (ch < 0 || ch >= npol) {
        std::cerr << "Requested polarity channel " << polarity << " is not available." << std::endl;
        return EXIT_FAILURE;
    }
    std::vector<double> x, y, z, startF, chirp;
    double adf = 1.0;
    const std::string pre = sandia ? "sph_MATData.Data." : "data.";
    const bool pos_ok = sandia ? per_pulse(find(v, pre + "radarCoordinateFrame.x"), Np, npol, ch, x) &&
                                     per_pulse(find(v, pre + "radarCoordinateFrame.y"), Np, npol, ch, y) &&
                                     per_pulse(find(v, pre + "radarCoordinateFrame.z"), Np, npol, ch, z)
                               : per_pulse(find(v, "data.x"), Np, 1, 0, x) &&
                                     per_pulse(find(v, "data.y"), Np, 1, 0, y) &&
                                     per_pulse(find(v, "data.z"), Np, 1, 0, z);
    if (!pos_ok) {
        std::cerr << "readData: not enough antenna positions for " << Np << " pulses" << std::endl;
        return EXIT_FAILURE;
    }
    if (sandia) {
        const Array *a = find(v, "sph_MATData.preamble.ADF");
        if (!a || !per_pulse(find(v, pre + "StartF"), Np, npol, ch, startF) ||
            !per_pulse(find(v, pre + "ChirpRate"), Np, npol, ch, chirp)) {
            std::cerr << "readData: Sandia record lacks ADF, StartF or ChirpRate" << std::endl;
            return EXIT_FAILURE;
        }
        adf = a->re[0];
    } else {
        const Array *f = find(v, "data.freq");
        if (!f || (f->numel() != (size_t)K && f->numel() != (size_t)K * Np) || K < 2) {
            std::cerr << "readData: GOTCHA record lacks data.freq (K or K x Np values)" << std::endl;
            return EXIT_FAILURE;
        }
        const bool each = f->numel() == (size_t)K * Np;
        for (int n = 0; n < Np; n++) {
            const double *fn = f->re.data() + (each ? (size_t)n * K : 0);
            startF.push_back(*std::min_element(fn, fn + K));
            chirp.push_back(std::fabs(fn[1] - fn[0]));
        }
    }
    const int numPulses = std::min(MAX_PULSES, Np);
    if (numPulses < 1) {
        std::cerr << "readData: MAX_PULSES = " << MAX_PULSES << std::endl;
        return EXIT_FAILURE;
    }
    const int first = (Np - numPulses) / 2;
    ph.sph_MATData_preamble_ADF = (float)adf;
    ph.numRangeSamples = K;
    ph.numAzimuthSamples = numPulses;
    ph.sph_MATData_Data_SampleData.resize(2 * (size_t)K * numPulses);
    ph.sph_MATData_Data_StartF.resize(numPulses);
    ph.sph_MATData_Data_ChirpRate.resize(numPulses);
    ph.sph_MATData_Data_radarCoordinateFrame_x.resize(numPulses);
    ph.sph_MATData_Data_radarCoordinateFrame_y.resize(numPulses);
    ph.sph_MATData_Data_radarCoordinateFrame_z.resize(numPulses);
    const size_t chan = (size_t)ch * K * Np;
    for (int i = 0; i < numPulses; i++) {
        const int n = first + i;
        for (int k = 0; k < K; k++) {
            const size_t s = chan + (size_t)n * K + k;
            ph.sph_MATData_Data_SampleData[2 * ((size_t)i * K + k)] = (float)sd->re[s];
            ph.sph_MATData_Data_SampleData[2 * ((size_t)i * K + k) + 1] = (float)sd->im[s];
        }
        ph.sph_MATData_Data_StartF[i] = (float)startF[n];
        ph.sph_MATData_Data_ChirpRate[i] = (float)(sandia ? chirp[n] : chirp[n] * adf);
        ph.sph_MATData_Data_radarCoordinateFrame_x[i] = (float)x[n];
        ph.sph_MATData_Data_radarCoordinateFrame_y[i] = (float)y[n];
        ph.sph_MATData_Data_radarCoordinateFrame_z[i] = (float)z[n];
    }
    if (verbose_)
        std::cout << "readData: " << (sandia ? "Sandia" : "GOTCHA") << " record, " << K
                  << " samples x " << Np << " pulses x " << npol << " channels; pulses " << first
                  << ".." << first + numPulses - 1 << " of channel " << ch << std::endl;
    return EXIT_SUCCESS;
}
