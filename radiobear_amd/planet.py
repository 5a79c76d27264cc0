"""Planet: the run entry point (radiobear/planet.py, radiobear/planet_base.py).

``Planet(name, config_file, ...)`` reads the configuration and builds the
atmosphere; ``run(freqs, b, scale, get_alpha, save_alpha, freqUnit, block,
reuse_override)`` computes the absorption table once on the GPU and then all
rays of the request in one batched launch (the reference loops over rays in
Python, planet.py:126-133).  The returned ``Data`` holds Tb as float32 like
the reference's.  Output files, logs and plots are not produced.
"""
import datetime
import os
from argparse import Namespace

import numpy as np

from . import set_utils, utils
from .alpha import Alpha
from .atmosphere import Atmosphere
from .brightness import Brightness
from .data_handling import Data


class Planet:
    planet_list = ['Jupiter', 'Saturn', 'Neptune', 'Uranus']

    def __init__(self, name, config_file='config.par', run_atm=True, load_formal=True, verbose=True,
                 precision='fp64', atmosphere=None, **kwargs):
        """precision: 'fp64' (plain) or 'fp32' (FP32 line sums); atmosphere: an
        Atmosphere or synth.SynthAtmosphere to use instead of the configured file."""
        from . import config as pcfg
        self.planet = name.capitalize()
        if self.planet not in self.planet_list:
            raise ValueError("{} not found.".format(self.planet))
        self.verbose = verbose
        self.load_formal = load_formal
        self.precision = precision
        cfile = config_file
        if cfile is not None and not os.path.exists(cfile):
            cand = os.path.join(self.planet, cfile)
            cfile = cand if os.path.exists(cand) else None
        self.config_file = cfile
        self.config = pcfg.planetConfig(self.planet, configFile=cfile)
        self.config.update_config(**kwargs)
        self.header = {}
        self.freqs = None
        self.freqUnit = None
        self.b = None
        self.scale = None
        self.get_alpha = None
        self.save_alpha = None
        self.data_type = None
        self.imSize = None
        self.rNorm = None
        self.data_return = Data()
        self.data_return.set('log', None)
        self.alpha_options = {'f': 'file', 'm': 'memory', 'n': 'none', 'c': 'none'}
        if atmosphere is not None:
            atm = atmosphere if isinstance(atmosphere, Atmosphere) else \
                Atmosphere.from_synthetic(atmosphere, self.planet, self.config)
            self.config = atm.config
            self.atmos = [atm]
        else:
            self.atmos = [Atmosphere(self.planet, idnum=i, config=self.config, verbose=verbose)
                          for i in range(len(self.config.gasFile))]
            if run_atm:
                for atm in self.atmos:
                    atm.std()
        self.alpha = [Alpha(idnum=i, config=self.config, load_formal=load_formal, verbose=verbose,
                            precision=precision) for i in range(len(self.atmos))]
        self.bright = Brightness(config=self.config, verbose=verbose)

    # --------------------------------------------------------------- requests
    def set_freqs(self, freqs, freqUnit='GHz'):
        self.header['freqs'] = '# freqs request: {} {}'.format(str(freqs), freqUnit)
        freqs, freqUnit = set_utils.set_freq(freqs, freqUnit)
        self.data_return.set('f', freqs)
        self.data_return.set('freqUnit', freqUnit)
        return freqs, freqUnit

    def set_b(self, b=(0.0, 0.0), block=(1, 1)):
        self.header['b'] = '# b request:  {}  {}'.format(str(b), str(block))
        rv = set_utils.set_b(b, block, Rpol=self.config.Rpol, Req=self.config.Req)
        self.b = rv.b
        self.block = rv.block
        self.data_type = rv.data_type
        self.imSize = rv.imSize
        self.data_return.set('b', self.b)

    def set_image(self):
        if self.data_type != 'image':
            return Namespace(true=False, block='_', imrow=[])
        if len(self.freqs) > 1:
            raise ValueError('Warning:  Image must be at only one frequency')
        block = '_'
        if abs(self.block[1]) > 1:
            block = '_{:02d}of{:02d}_'.format(self.block[0], abs(self.block[1]))
        return Namespace(true=True, i=None, block=block, imrow=[])

    def check_reuse(self, freqs, scale, get_alpha, save_alpha, reuse_override='check'):
        """Whether the last absorption table can be reused (planet_base.py:303-346)."""
        if reuse_override == 'true':
            return True
        if reuse_override == 'false':
            return False
        if get_alpha != self.get_alpha or save_alpha != self.save_alpha:
            return False
        if self.freqs is None or len(freqs) != len(self.freqs):
            return False
        for a, b in zip(sorted(self.freqs), sorted(freqs)):
            if (a - b) / a > 0.0001:
                return False
        if utils.isanynum(scale) and utils.isanynum(self.scale):
            if (scale - self.scale) / scale > 0.0001:
                return False
        if not isinstance(scale, type(self.scale)):
            return False
        if isinstance(scale, (list, np.ndarray)):
            if len(scale) != len(self.scale):
                return False
            for a, b in zip(scale, self.scale):
                if (a - b) / a > 0.0001:
                    return False
        if isinstance(scale, dict):
            if sorted(scale) != sorted(self.scale):
                return False
            for k in scale:
                if len(scale[k]) != len(self.scale[k]):
                    return False
                for a, b in zip(scale[k], self.scale[k]):
                    if (a - b) / a > 0.001:
                        return False
        return True

    def alpha_layers(self, freqs, atmos, scale=False, get_alpha='calc', save_alpha='none'):
        for i, atm in enumerate(atmos):
            self.alpha[i].reset_layers()
            self.alpha[i].get_layers(freqs=freqs, atm=atm, scale=scale, get_alpha=get_alpha,
                                     save_alpha=save_alpha)

    # -------------------------------------------------------------------- run
    def run(self, freqs, b='disc', scale=False, get_alpha='calc', save_alpha='none', freqUnit='GHz',
            block=(1, 1), reuse_override='check'):
        """Brightness temperatures for a frequency and b request (planet.py:59-160)."""
        get_alpha = self.alpha_options[get_alpha[0].lower()]
        save_alpha = self.alpha_options[save_alpha[0].lower()]
        freqs, freqUnit = self.set_freqs(freqs, freqUnit)
        reuse = self.check_reuse(freqs, scale, get_alpha, save_alpha, reuse_override=reuse_override.lower())
        self.set_b(b, block)
        # the ray paths do not depend on the absorption: launched first, they run beside it
        paths = self.bright.start_paths(self.b, self.atmos[0], self.config.orientation)
        if not reuse:
            self.freqs = freqs
            self.freqUnit = utils.proc_unit(freqUnit)
            self.scale = scale
            self.get_alpha = get_alpha
            self.save_alpha = save_alpha
            self.alpha_layers(self.freqs, self.atmos, scale=scale, get_alpha=get_alpha, save_alpha=save_alpha)
        is_img = self.set_image()
        run_start = datetime.datetime.now()
        Tb = self.bright.batch(self.b, self.freqs, self.atmos[0], self.alpha[0], self.config.orientation,
                               paths=paths)
        run_stop = datetime.datetime.now()
        if is_img.true:
            n = self.imSize[0]
            self.Tb = [list(Tb[i:i + n, 0]) for i in range(0, len(Tb), n)]
        else:
            self.Tb = [list(row) for row in Tb]
        on = self.bright.nsteps >= 0
        cfg = self.config
        if on.any():
            self.rNorm = self.atmos[0].property[cfg.LP['R']][0]
            tip, rotate = _aspect(cfg)
            self.header['orientation'] = '# orientation:   {}'.format(repr(cfg.orientation))
            self.header['aspect'] = '# aspect tip, rotate:  {:.4f}  {:.4f}'.format(utils.r2d(tip), utils.r2d(rotate))
            self.header['rNorm'] = '# rNorm: {}'.format(self.rNorm)
            if self.data_type == 'image':                  # planet_base.py:370-375
                self.header['imgSize'] = '# imgSize: {}'.format(self.imSize)
                resolution = utils.r2asec(np.arctan(abs(self.b[1][0] - self.b[0][0]) * self.rNorm / cfg.distance))
                self.header['res'] = '# res:  {} arcsec'.format(resolution)
        else:
            self.rNorm = None
            self.header['orientation'] = '# orientation not set'
            self.header['aspect'] = '# aspect tip, rotate not set'
            self.header['rNorm'] = '# rNorm not set'
        self.header['data-type'] = '#* type:  {}'.format(self.data_type)
        self.header['gtype'] = '# gtype: {}'.format(cfg.gtype)
        self.header['radii'] = '# radii:  {:.1f}  {:.1f}  km'.format(cfg.Req, cfg.Rpol)
        self.header['distance'] = '# distance:  {} km'.format(cfg.distance)
        self.header['start'] = "#* start: {:%Y-%m-%d %H:%M:%S}".format(run_start)
        self.header['stop'] = "#* stop: {:%Y-%m-%d %H:%M:%S}".format(run_stop)
        self.data_return.set('start', run_start)
        self.data_return.set('stop', run_stop)
        self.data_return.set('Tb', self.Tb)
        self.data_return.set('type', self.data_type)
        self.data_return.set('header', self.header)
        if cfg.write_output_files:                          # planet.py:139-147
            from .fileio import FileIO, output_name
            out = output_name(self.planet, self.data_type, is_img.block, run_start, cfg.output_directory)
            FileIO(cfg.output_directory, os.path.join(cfg.scratch_directory, 'specoutputline.dat')).write(
                out, self.data_return)
            self.output_file = out
        return self.data_return


def _aspect(cfg):
    from .raypath import computeAspect
    return computeAspect(cfg.orientation, 1.0 - cfg.Rpol / cfg.Req)
