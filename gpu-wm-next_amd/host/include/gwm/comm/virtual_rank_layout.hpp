// Reference include path include/gwm/comm/virtual_rank_layout.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
