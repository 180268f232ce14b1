// Reference include path include/gwm/domain/subdomain_descriptor.hpp, served by the host mirror.
#pragma once
#include "gwm/gwm_dycore.hpp"
