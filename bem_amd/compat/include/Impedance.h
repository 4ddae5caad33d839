// Drop-in for the reference's src/Impedance.h: Zts and Zdense with the
// reference's signatures (src/Impedance.h:20-21) over the C ABI.
#pragma once

#include "../bem_compat.h"

using namespace std;
