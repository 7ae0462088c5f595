// Generated by laughlinedge_amd/system_parameters.cpp
#ifndef SYSPARAMS_H
#define SYSPARAMS_H

namespace sys_params {
    constexpr int particlesNumber = 25;
    constexpr int inverseFilling = 2;

    constexpr int confinementExponent = 4;
    constexpr float confinementStrength = 1.000000;
    constexpr double confinementW = 0.000000000000;
    constexpr double confinementEps = 0.000000000000;

    constexpr float RCl = 10.000000;
    constexpr float DCl = 20.000000;
    constexpr float RCl_reciprocal = 0.100000;
    constexpr float R0_reciprocal = 0.200000;

    constexpr int ansatzExtendedMaxDegree = 13;
    constexpr int ansatzMaxPartitionSize = 5;
    constexpr int ansatzSubspaceDimension = 80;

    constexpr int ansatzMultipletSize = 5;

    constexpr int ansatzMaxAngularMomentum = 15;
    constexpr int angularMomentumCutoff = 63;

}

#endif
