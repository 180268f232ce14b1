// Reference include path include/gwm/dycore/passive_tracer.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
