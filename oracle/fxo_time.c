/* Newmark / CentralDifference restatement. */
#include "fxo.h"
