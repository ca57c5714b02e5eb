// Space.h -- see Structure.h (the reference declares Space in its own header).
#pragma once
#include "Structure.h"
