// Reference include path include/gwm/core/moist_thermo.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_io.hpp"
