// Reference include path include/gwm/dycore/boundary_conditions.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
