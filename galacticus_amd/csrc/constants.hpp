// constants.hpp -- physical constants and unit conversions of the engine.
//
// Definitions follow source/numerical/constants/{astronomical,atomic,
// physical,units,prefixes}.F90 of the reference; numerical values are
// CODATA 2018 / IAU 2015 where the reference takes them from external
// tables.  Units of the ODE state: Msun, Mpc, km/s, Gyr.
#pragma once

namespace gal {

constexpr double kPi = 3.14159265358979323846;
constexpr double kKilo = 1.0e3;
constexpr double kHecto = 1.0e2;
constexpr double kCenti = 1.0e-2;
constexpr double kParsec = 3.0856775814913673e16;           // m
constexpr double kMegaParsec = 1.0e6 * kParsec;             // m
constexpr double kYear = 3.15581497635456e7;                // s (astronomical.F90 "year")
constexpr double kGigaYear = 1.0e9 * kYear;                 // s
constexpr double kMassSolar = 1.98847e30;                   // kg
constexpr double kGravitySI = 6.67430e-11;                  // m^3 kg^-1 s^-2
// gravitationalConstant_internal: Mpc (km/s)^2 / Msun
constexpr double kGravity = kGravitySI * kMassSolar / (kKilo * kKilo) / kMegaParsec;
// MpcPerKmPerSToGyr
constexpr double kMpcPerKmsToGyr = kMegaParsec / kKilo / kGigaYear;
constexpr double kAtomicMassUnit = 1.66053906660e-27;       // kg
constexpr double kAtomicMassHydrogen = 1.0078250322;
constexpr double kAtomicMassHelium = 4.002602;
constexpr double kMassHydrogenAtom = kAtomicMassHydrogen * kAtomicMassUnit;
constexpr double kBoltzmann = 1.380649e-23;                 // J/K
constexpr double kSpeedLight = 2.99792458e8;                // m/s
constexpr double kThomson = 6.6524587321e-29;               // m^2
constexpr double kErgs = 1.0e-7;                            // J
constexpr double kHydrogenPrimordial = 0.7514;
constexpr double kHeliumPrimordial = 0.2486;
constexpr double kMetallicitySolar = 0.0188;
// meanAtomicMassPrimordial = 1/(2X/A_H + 3Y/A_He)
constexpr double kMeanAtomicMassPrimordial =
    1.0 / (2.0 * kHydrogenPrimordial / kAtomicMassHydrogen + 3.0 * kHeliumPrimordial / kAtomicMassHelium);

}  // namespace gal
