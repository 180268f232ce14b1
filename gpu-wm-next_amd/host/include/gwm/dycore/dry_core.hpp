// Reference include path include/gwm/dycore/dry_core.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
