// sarf_cxx_api.cpp -- definitions the C++ callback interface
// (include/charlotte_sar_api.hpp) needs from the library.
#include <iostream>
#include <string>

#include "charlotte_sar_api.hpp"

// Output file of the sph_sar_data_callback_* functions and the verbosity
// flag (charlotte_sar_api.hpp:32-33 of the reference).
std::string outputfile = "sar_image.bmp";
bool verbose = false;

int readData(const std::string &inputfile, const int MAX_PULSES, const std::string &polarity,
             PhaseHistory<float> &ph, bool verbose_) {
    (void)MAX_PULSES;
    (void)polarity;
    (void)ph;
    (void)verbose_;
    std::cerr << "readData(" << inputfile
              << "): MAT-file input is not provided by libsarf; build a PhaseHistory from "
                 "arrays and call sph_sar_data_callback_gpu" << std::endl;
    return EXIT_FAILURE;
}
