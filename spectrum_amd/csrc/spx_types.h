// Record layouts shared by host and device code come from the public C ABI header.
#pragma once
#include <spectrum_mi355x.h>
