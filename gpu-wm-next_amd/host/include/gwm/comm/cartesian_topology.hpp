// Reference include path include/gwm/comm/cartesian_topology.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
