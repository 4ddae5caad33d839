// Drop-in for the reference's src/HeatTransfer.h: a driver that includes it
// (src/transfer.cpp:1) compiles unchanged and links libbem_compat.so and
// libbem_hip.so.  CalcHeatTransfer keeps the reference's signature
// (src/HeatTransfer.h:33).  parseShapes (src/HeatTransfer.h:35) takes the
// tree pointers by reference: the reference passes them by value, so its
// callers never received the trees; a caller written against the reference
// passes the same pointer variables.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "../bem_compat.h"
#include "FrequencySweep.h"
#include "Impedance.h"
#include "ReferenceCountingPointer.h"

using namespace std;
