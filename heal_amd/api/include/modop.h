// modop.h — part of the HEAL hardware API on MI355X; all declarations live in
// lattica_hw_api.h (this name is kept for callers that include it directly).
#ifndef HEAL_AMD_MODOP_H
#define HEAL_AMD_MODOP_H
#include "lattica_hw_api.h"
#endif
