// Reference include path include/gwm/domain/grid_metrics.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
