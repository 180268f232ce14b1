// Reference include path include/gwm/dycore/idealized_cases.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
