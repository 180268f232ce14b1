// Reference include path include/gwm/dycore/dry_diagnostics.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
