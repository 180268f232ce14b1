// Reference include path include/gwm/state/field3d.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
