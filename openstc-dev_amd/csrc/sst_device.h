// k-omega SST model constants (Menter's SST: inner set 1, k-epsilon set 2;
// SST.cuh:13-34), shared by the source kernel, the viscous face flux and the
// wall boundary condition.
#pragma once

namespace ostc {
namespace sst {
constexpr double kBetaStar = 0.09;
constexpr double kSqrtBetaStar = 0.3;
constexpr double kKappa = 0.41;
constexpr double kSigmaK1 = 0.85;
constexpr double kSigmaOmega1 = 0.5;
constexpr double kBeta1 = 0.0750;
constexpr double kA1 = 0.31;
constexpr double kSigmaK2 = 1.0;
constexpr double kSigmaOmega2 = 0.856;
constexpr double kBeta2 = 0.0828;
constexpr double kGamma1 = kBeta1 / kBetaStar - kSigmaOmega1 * kKappa * kKappa / kSqrtBetaStar;
constexpr double kGamma2 = kBeta2 / kBetaStar - kSigmaOmega2 * kKappa * kKappa / kSqrtBetaStar;
}  // namespace sst
}  // namespace ostc
