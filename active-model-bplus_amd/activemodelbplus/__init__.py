"""MI355X-native Active Model B+ (conserved active scalar field) integrator.

Only the ``activemodelbplus.integrator`` subpackage is provided; the reference's
mean-field theory modules (activefield, activedroplets, ode, ...) are not part of
this package.
"""
