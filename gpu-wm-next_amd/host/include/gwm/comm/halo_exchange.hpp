// Reference include path include/gwm/comm/halo_exchange.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
