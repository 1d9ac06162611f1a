"""Actions shared by GUI bridges (``xentica/bridge/base.py:14-39``)."""

__all__ = ['Bridge', ]


class Bridge:
    """Model actions a GUI can trigger; each is ``func(model, gui)``."""

    @staticmethod
    def noop(_model, _gui):
        """Do nothing."""

    @staticmethod
    def exit_app(_model, gui):
        """Ask the GUI to exit."""
        gui.exit_app()

    @staticmethod
    def speed(dspeed):
        """Action changing the simulation speed by ``dspeed``."""
        def func(model, _gui):
            model.apply_speed(dspeed)
        return func

    @staticmethod
    def toggle_pause(model, _gui):
        """Pause or resume the model."""
        model.toggle_pause()

    @staticmethod
    def toggle_sysinfo(_model, gui):
        """Show or hide the GUI's system info."""
        gui.sysinfo.toggle()
