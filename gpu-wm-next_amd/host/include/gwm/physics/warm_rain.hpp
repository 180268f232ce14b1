// Reference include path include/gwm/physics/warm_rain.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
