// Forwarding header: the reference include/alto.hpp name, so that code written
// against the reference (for example its src/main.cpp) builds unchanged with
// -I include/alto/ref and links libalto_hip.
#pragma once
#include "../alto.hpp"
