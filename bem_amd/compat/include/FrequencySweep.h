// Drop-in for the reference's src/FrequencySweep.h: FrequencySweep(char *)
// reading one frequency per line into f (src/FrequencySweep.h:9-13).
#pragma once

#include <vector>

#include "../bem_compat.h"

using namespace std;
