#pragma once
#include <spectrum/types/spectrum_types.hpp>
