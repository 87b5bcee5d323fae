// jk_common.h -- J/K digestion and dense-scatter kernel templates over the
// stored ERI classes.
#pragma once
#include "eri_common.h"

