// Drop-in for the reference's src/util/ReferenceCountingPointer.h as far as a
// driver of the hot path sees it: the initialisation call every reference
// driver makes first (src/transfer.cpp:5).  The reference needs it before any
// matrix copy (src/util/ReferenceCountingPointer.h:58-61); here the matrices
// live on the device behind the C ABI, so it only sets up the process-wide
// device context early.
#pragma once

#include "../bem_compat.h"   // declares initReferenceCountingPointerLock()
