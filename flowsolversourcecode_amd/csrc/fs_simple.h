// Internal include of the SIMPLE C ABI.
#pragma once
#include "fsamd_simple.h"
